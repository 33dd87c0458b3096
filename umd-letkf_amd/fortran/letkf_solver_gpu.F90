!> @file letkf_solver_gpu.F90
!! Fortran host of the GPU local analysis: the init / run / final triple of
!! the reference's letkf_solver module (src/letkf_solver.F90:36-38), with
!! the reference's module arrays passed explicitly.
!!
!! letkf_solver_gpu_init  replaces letkf_solver_init's solver set-up
!!                        (src/letkf_solver.F90:73-142): problem sizes,
!!                        precision/path, localizer, state variables,
!!                        inflation
!! letkf_solver_gpu_run   replaces the body of letkf_solver_run
!!                        (src/letkf_solver.F90:150-389): uploads the
!!                        background and the observations, runs the local
!!                        analysis and writes the analysis back into
!!                        state_ij / state_mean_ij / state_sprd_ij in place,
!!                        like the reference
!! letkf_solver_gpu_final releases the device resources
!!
!! Every routine returns the library status in ierr (LETKF_OK = 0); the
!! reference's convention for a failure is letkf_mpi_abort(msg) with
!! msg = letkf_gpu_errmsg(letkf_solver_gpu_handle()).
MODULE letkf_solver_gpu
  USE, INTRINSIC :: iso_c_binding
  USE letkf_gpu_mod
  IMPLICIT NONE
  PRIVATE

  PUBLIC :: letkf_solver_gpu_init, letkf_solver_gpu_run, letkf_solver_gpu_final
  PUBLIC :: letkf_solver_gpu_handle, letkf_solver_gpu_diag

  TYPE(c_ptr), SAVE :: handle = c_null_ptr
  INTEGER(c_int64_t), SAVE :: n_k = 0, n_ns = 0, n_ij = 0

CONTAINS

  FUNCTION letkf_solver_gpu_handle() RESULT(h)
    TYPE(c_ptr) :: h
    h = handle
  END FUNCTION letkf_solver_gpu_handle

  !> ens_size, grid_ns, ij_count: the reference's sizes; loc: the localizer
  !! descriptor; statevars as (grid_s_idx 0-based, levels, ana_inc_max,
  !! ana_bounds(2, nvars)); rtps, rtpp, mul: solver.inflation
  SUBROUTINE letkf_solver_gpu_init(ens_size, grid_ns, ij_count, precision, path, loc, &
       grid_s_idx, levels, ana_inc_max, ana_bounds, rtps, rtpp, mul, ierr)
    INTEGER, INTENT(in) :: ens_size, grid_ns, ij_count
    INTEGER(c_int32_t), INTENT(in) :: precision, path
    TYPE(letkf_loc_desc), INTENT(in) :: loc
    INTEGER(c_int32_t), INTENT(in) :: grid_s_idx(:), levels(:)
    REAL(c_float), INTENT(in) :: ana_inc_max(:), ana_bounds(:,:)
    REAL, INTENT(in) :: rtps, rtpp, mul
    INTEGER, INTENT(out) :: ierr
    TYPE(letkf_gpu_dims) :: dims

    dims%ens_size = ens_size
    dims%grid_ns = grid_ns
    dims%ij_count = ij_count
    dims%max_obs = 0
    dims%precision = precision
    dims%path = path
    ierr = letkf_gpu_create(dims, handle)
    IF (ierr /= LETKF_OK) RETURN
    n_k = ens_size
    n_ns = grid_ns
    n_ij = ij_count
    ierr = letkf_gpu_set_statevars(handle, INT(SIZE(levels), c_int32_t), grid_s_idx, levels, &
         ana_inc_max, ana_bounds)
    IF (ierr /= LETKF_OK) RETURN
    ierr = letkf_gpu_set_localizer(handle, loc)
    IF (ierr /= LETKF_OK) RETURN
    ierr = letkf_gpu_set_solver(handle, REAL(rtps, c_float), REAL(rtpp, c_float), &
         REAL(mul, c_float))
  END SUBROUTINE letkf_solver_gpu_init

  !> state_ij(k, ns, ij) background perturbations in, analysis out (with the
  !! mean added back, masked points untouched); mean and sprd (ns, ij) in/out.
  !! Observations: the QC-compacted obs_def fields, obs_hx(k, nobs)
  !! perturbations and obs_hx_mean(nobs).
  SUBROUTINE letkf_solver_gpu_run(state_ij, state_mean_ij, state_sprd_ij, lat_ij, lon_ij, &
       mask_ij, olat, olon, ozdim, otime, oval, oerr, obsid, platid, obs_hx, obs_hx_mean, ierr)
    REAL(c_float), INTENT(inout) :: state_ij(:,:,:), state_mean_ij(:,:), state_sprd_ij(:,:)
    REAL(c_float), INTENT(in) :: lat_ij(:), lon_ij(:)
    LOGICAL, INTENT(in) :: mask_ij(:)
    REAL(c_float), INTENT(in) :: olat(:), olon(:), ozdim(:), otime(:), oval(:), oerr(:)
    INTEGER(c_int32_t), INTENT(in) :: obsid(:), platid(:)
    REAL(c_float), INTENT(in) :: obs_hx(:,:), obs_hx_mean(:)
    INTEGER, INTENT(out) :: ierr
    INTEGER(c_int32_t), ALLOCATABLE :: imask(:)

    ALLOCATE(imask(SIZE(mask_ij)))
    imask = MERGE(1_c_int32_t, 0_c_int32_t, mask_ij)
    ierr = letkf_gpu_set_state(handle, state_ij, state_mean_ij, state_sprd_ij, lat_ij, &
         lon_ij, imask)
    IF (ierr /= LETKF_OK) RETURN
    ierr = letkf_gpu_set_obs(handle, INT(SIZE(olat), c_int64_t), olat, olon, ozdim, otime, &
         oval, oerr, obsid, platid, obs_hx, obs_hx_mean)
    IF (ierr /= LETKF_OK) RETURN
    ierr = letkf_gpu_run(handle)
    IF (ierr /= LETKF_OK) RETURN
    ierr = letkf_gpu_get_state(handle, state_ij, state_mean_ij, state_sprd_ij)
  END SUBROUTINE letkf_solver_gpu_run

  !> diagnostics of the last run (diag.solver.nc fields)
  SUBROUTINE letkf_solver_gpu_diag(maxhz, obscount, lg_obscount, lg_obsloc, ierr)
    REAL(c_float), INTENT(out), TARGET :: maxhz(:), obscount(:), lg_obscount(:,:), lg_obsloc(:,:)
    INTEGER, INTENT(out) :: ierr
    TYPE(letkf_gpu_diag) :: d
    d%maxhz = c_loc(maxhz)
    d%obscount = c_loc(obscount)
    d%lg_obscount = c_loc(lg_obscount)
    d%lg_obsloc = c_loc(lg_obsloc)
    ierr = letkf_gpu_get_diag(handle, d)
  END SUBROUTINE letkf_solver_gpu_diag

  SUBROUTINE letkf_solver_gpu_final(ierr)
    INTEGER, INTENT(out) :: ierr
    ierr = letkf_gpu_destroy(handle)
    handle = c_null_ptr
  END SUBROUTINE letkf_solver_gpu_final

END MODULE letkf_solver_gpu
