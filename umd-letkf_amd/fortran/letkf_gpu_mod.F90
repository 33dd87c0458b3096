!> @file letkf_gpu_mod.F90
!! iso_c_binding interface to libletkf_gpu.so (include/letkf_gpu.h).
!!
!! The derived types are BIND(C) images of the C structs, so they can be
!! filled from the reference's module variables and passed by reference.
!! Every function returns LETKF_OK (0) or a LETKF_E* status;
!! letkf_gpu_errmsg() turns letkf_gpu_last_error() into a Fortran string.
!! Arrays keep the reference's column-major layout; indices at this
!! interface are 0-based (gridpoint, slab, group, obs).
MODULE letkf_gpu_mod
  USE, INTRINSIC :: iso_c_binding
  IMPLICIT NONE
  PRIVATE

  ! status codes
  INTEGER(c_int), PARAMETER, PUBLIC :: LETKF_OK = 0, LETKF_EINVAL = 1, &
       LETKF_EUNSUPPORTED = 2, LETKF_ENOMEM = 3, LETKF_EHIP = 4, &
       LETKF_ECAPACITY = 5, LETKF_ESTATE = 6
  ! arithmetic and execution path
  INTEGER(c_int32_t), PARAMETER, PUBLIC :: LETKF_PREC_FP64 = 0, LETKF_PREC_FP32 = 1, &
       LETKF_PREC_BF16X3 = 2, LETKF_PREC_BF16 = 3, LETKF_PREC_BF16X4 = 4
  INTEGER(c_int32_t), PARAMETER, PUBLIC :: LETKF_PATH_PLAIN = 0, LETKF_PATH_BATCHED = 1
  ! localizer classes (localization.class)
  INTEGER(c_int32_t), PARAMETER, PUBLIC :: LETKF_LOC_NOVRT = 1, LETKF_LOC_OCEAN = 2, &
       LETKF_LOC_VGAUSS = 3
  INTEGER, PARAMETER, PUBLIC :: LETKF_MAX_INTERP = 32, LETKF_MAX_IDLIST = 64

  !> letkf_gpu_dims
  TYPE, BIND(C), PUBLIC :: letkf_gpu_dims
     INTEGER(c_int64_t) :: ens_size = 0
     INTEGER(c_int64_t) :: grid_ns = 0
     INTEGER(c_int64_t) :: ij_count = 0
     INTEGER(c_int64_t) :: max_obs = 0
     INTEGER(c_int32_t) :: precision = 0
     INTEGER(c_int32_t) :: path = 0
  END TYPE letkf_gpu_dims

  !> letkf_linearinterp_lat: the YAML value list of a linearinterp_lat table
  TYPE, BIND(C), PUBLIC :: letkf_linearinterp_lat
     INTEGER(c_int32_t) :: npoints = 0
     REAL(c_float) :: lat(32) = 0.0
     REAL(c_float) :: radius(32) = 0.0
  END TYPE letkf_linearinterp_lat

  !> letkf_loc_desc
  TYPE, BIND(C), PUBLIC :: letkf_loc_desc
     INTEGER(c_int32_t) :: cls = 0
     TYPE(letkf_linearinterp_lat) :: hzloc
     TYPE(letkf_linearinterp_lat) :: hzloc_prof
     TYPE(letkf_linearinterp_lat) :: hzloc_sat
     REAL(c_float) :: tloc_prof = -1.0
     REAL(c_float) :: tloc_sat = -1.0
     INTEGER(c_int32_t) :: vtloc_surf = 0
     REAL(c_float) :: bkg_t_delta = 0.0
     INTEGER(c_int32_t) :: bkg_t_var_idx = 0
     INTEGER(c_int32_t) :: bkg_t_var_len = 0
     INTEGER(c_int32_t) :: n_sat_obs = 0
     INTEGER(c_int32_t) :: sat_obs(64) = 0
     INTEGER(c_int32_t) :: n_sat_plats = 0
     INTEGER(c_int32_t) :: sat_plats(64) = 0
     REAL(c_float) :: vt_scale = 0.0
     INTEGER(c_int32_t) :: nlev = 0
     TYPE(c_ptr) :: lev_depth = c_null_ptr
  END TYPE letkf_loc_desc

  !> letkf_gpu_diag: any pointer may be c_null_ptr
  TYPE, BIND(C), PUBLIC :: letkf_gpu_diag
     TYPE(c_ptr) :: maxhz = c_null_ptr
     TYPE(c_ptr) :: obscount = c_null_ptr
     TYPE(c_ptr) :: lg_obscount = c_null_ptr
     TYPE(c_ptr) :: lg_obsloc = c_null_ptr
     TYPE(c_ptr) :: lg_selhash = c_null_ptr
  END TYPE letkf_gpu_diag

  !> letkf_gpu_timing
  TYPE, BIND(C), PUBLIC :: letkf_gpu_timing
     REAL(c_double) :: search_ms, product_ms, eigen_ms, finalize_ms, total_ms
     INTEGER(c_int64_t) :: n_problems, n_pairs, n_cand, n_launch
     REAL(c_double) :: order_ms
  END TYPE letkf_gpu_timing

  !> letkf_gpu_tuning
  TYPE, BIND(C), PUBLIC :: letkf_gpu_tuning
     INTEGER(c_int32_t) :: order, bucket_bits, tridiag_reg, n_tried
     REAL(c_double) :: best_ms, default_ms
     INTEGER(c_int32_t) :: overlap, chunk_log2, xcd, tile
  END TYPE letkf_gpu_tuning

  PUBLIC :: letkf_gpu_version, letkf_gpu_create, letkf_gpu_destroy, letkf_gpu_last_error, &
       letkf_gpu_set_state, letkf_gpu_set_obs, letkf_gpu_set_statevars, &
       letkf_gpu_set_localizer, letkf_gpu_set_solver, letkf_gpu_run, letkf_gpu_get_state, &
       letkf_gpu_get_state_f64, letkf_gpu_get_diag, letkf_gpu_maxgroups, &
       letkf_gpu_get_selection, letkf_gpu_get_timing, letkf_gpu_errmsg, &
       letkf_gpu_set_state_members, letkf_gpu_set_obs_members, letkf_gpu_get_state_members, &
       letkf_gpu_autotune, letkf_gpu_get_tuning, letkf_gpu_set_tuning

  INTERFACE
     FUNCTION letkf_gpu_version() BIND(C, name="letkf_gpu_version")
       IMPORT :: c_ptr
       TYPE(c_ptr) :: letkf_gpu_version
     END FUNCTION letkf_gpu_version

     FUNCTION letkf_gpu_create(dims, handle) BIND(C, name="letkf_gpu_create")
       IMPORT :: c_int, c_ptr, letkf_gpu_dims
       TYPE(letkf_gpu_dims), INTENT(in) :: dims
       TYPE(c_ptr), INTENT(out) :: handle
       INTEGER(c_int) :: letkf_gpu_create
     END FUNCTION letkf_gpu_create

     FUNCTION letkf_gpu_destroy(h) BIND(C, name="letkf_gpu_destroy")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int) :: letkf_gpu_destroy
     END FUNCTION letkf_gpu_destroy

     FUNCTION letkf_gpu_last_error(h) BIND(C, name="letkf_gpu_last_error")
       IMPORT :: c_ptr
       TYPE(c_ptr), VALUE :: h
       TYPE(c_ptr) :: letkf_gpu_last_error
     END FUNCTION letkf_gpu_last_error

     FUNCTION letkf_gpu_set_state(h, state_ij, mean_ij, sprd_ij, lat_ij, lon_ij, mask_ij) &
          BIND(C, name="letkf_gpu_set_state")
       IMPORT :: c_int, c_ptr, c_float, c_int32_t
       TYPE(c_ptr), VALUE :: h
       REAL(c_float), INTENT(in) :: state_ij(*), mean_ij(*), sprd_ij(*), lat_ij(*), lon_ij(*)
       INTEGER(c_int32_t), INTENT(in) :: mask_ij(*)
       INTEGER(c_int) :: letkf_gpu_set_state
     END FUNCTION letkf_gpu_set_state

     FUNCTION letkf_gpu_set_obs(h, n_obs, lat, lon, zdim, time, val, err, obsid, platid, &
          hx_pert, hx_mean) BIND(C, name="letkf_gpu_set_obs")
       IMPORT :: c_int, c_ptr, c_float, c_int32_t, c_int64_t
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int64_t), VALUE :: n_obs
       REAL(c_float), INTENT(in) :: lat(*), lon(*), zdim(*), time(*), val(*), err(*)
       INTEGER(c_int32_t), INTENT(in) :: obsid(*), platid(*)
       REAL(c_float), INTENT(in) :: hx_pert(*), hx_mean(*)
       INTEGER(c_int) :: letkf_gpu_set_obs
     END FUNCTION letkf_gpu_set_obs

     FUNCTION letkf_gpu_set_statevars(h, nvars, grid_s_idx, levels, ana_inc_max, ana_bounds) &
          BIND(C, name="letkf_gpu_set_statevars")
       IMPORT :: c_int, c_ptr, c_float, c_int32_t
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int32_t), VALUE :: nvars
       INTEGER(c_int32_t), INTENT(in) :: grid_s_idx(*), levels(*)
       REAL(c_float), INTENT(in) :: ana_inc_max(*), ana_bounds(*)
       INTEGER(c_int) :: letkf_gpu_set_statevars
     END FUNCTION letkf_gpu_set_statevars

     FUNCTION letkf_gpu_set_localizer(h, desc) BIND(C, name="letkf_gpu_set_localizer")
       IMPORT :: c_int, c_ptr, letkf_loc_desc
       TYPE(c_ptr), VALUE :: h
       TYPE(letkf_loc_desc), INTENT(in) :: desc
       INTEGER(c_int) :: letkf_gpu_set_localizer
     END FUNCTION letkf_gpu_set_localizer

     FUNCTION letkf_gpu_set_solver(h, rtps, rtpp, mul) BIND(C, name="letkf_gpu_set_solver")
       IMPORT :: c_int, c_ptr, c_float
       TYPE(c_ptr), VALUE :: h
       REAL(c_float), VALUE :: rtps, rtpp, mul
       INTEGER(c_int) :: letkf_gpu_set_solver
     END FUNCTION letkf_gpu_set_solver

     FUNCTION letkf_gpu_run(h) BIND(C, name="letkf_gpu_run")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int) :: letkf_gpu_run
     END FUNCTION letkf_gpu_run

     FUNCTION letkf_gpu_get_state(h, state_ij, mean_ij, sprd_ij) &
          BIND(C, name="letkf_gpu_get_state")
       IMPORT :: c_int, c_ptr, c_float
       TYPE(c_ptr), VALUE :: h
       REAL(c_float), INTENT(out) :: state_ij(*), mean_ij(*), sprd_ij(*)
       INTEGER(c_int) :: letkf_gpu_get_state
     END FUNCTION letkf_gpu_get_state

     FUNCTION letkf_gpu_get_state_f64(h, state_ij, mean_ij, sprd_ij) &
          BIND(C, name="letkf_gpu_get_state_f64")
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h
       REAL(c_double), INTENT(out) :: state_ij(*), mean_ij(*), sprd_ij(*)
       INTEGER(c_int) :: letkf_gpu_get_state_f64
     END FUNCTION letkf_gpu_get_state_f64

     FUNCTION letkf_gpu_get_diag(h, diag) BIND(C, name="letkf_gpu_get_diag")
       IMPORT :: c_int, c_ptr, letkf_gpu_diag
       TYPE(c_ptr), VALUE :: h
       TYPE(letkf_gpu_diag), INTENT(in) :: diag
       INTEGER(c_int) :: letkf_gpu_get_diag
     END FUNCTION letkf_gpu_get_diag

     FUNCTION letkf_gpu_maxgroups(h, maxgroups) BIND(C, name="letkf_gpu_maxgroups")
       IMPORT :: c_int, c_ptr, c_int32_t
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int32_t), INTENT(out) :: maxgroups
       INTEGER(c_int) :: letkf_gpu_maxgroups
     END FUNCTION letkf_gpu_maxgroups

     FUNCTION letkf_gpu_get_selection(h, ij, group, idx_out, rloc_out, cap, m_out) &
          BIND(C, name="letkf_gpu_get_selection")
       IMPORT :: c_int, c_ptr, c_int32_t, c_int64_t, c_float
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int64_t), VALUE :: ij
       INTEGER(c_int32_t), VALUE :: group
       INTEGER(c_int32_t), INTENT(out) :: idx_out(*)
       REAL(c_float), INTENT(out) :: rloc_out(*)
       INTEGER(c_int64_t), VALUE :: cap
       INTEGER(c_int64_t), INTENT(out) :: m_out
       INTEGER(c_int) :: letkf_gpu_get_selection
     END FUNCTION letkf_gpu_get_selection

     FUNCTION letkf_gpu_get_timing(h, t) BIND(C, name="letkf_gpu_get_timing")
       IMPORT :: c_int, c_ptr, letkf_gpu_timing
       TYPE(c_ptr), VALUE :: h
       TYPE(letkf_gpu_timing), INTENT(out) :: t
       INTEGER(c_int) :: letkf_gpu_get_timing
     END FUNCTION letkf_gpu_get_timing

     FUNCTION letkf_gpu_set_state_members(h, members, lat_ij, lon_ij, mask_ij) &
          BIND(C, name="letkf_gpu_set_state_members")
       IMPORT :: c_int, c_ptr, c_float, c_int32_t
       TYPE(c_ptr), VALUE :: h
       REAL(c_float), INTENT(in) :: members(*), lat_ij(*), lon_ij(*)
       INTEGER(c_int32_t), INTENT(in) :: mask_ij(*)
       INTEGER(c_int) :: letkf_gpu_set_state_members
     END FUNCTION letkf_gpu_set_state_members

     FUNCTION letkf_gpu_set_obs_members(h, n_obs, lat, lon, zdim, time, val, err, obsid, &
          platid, qc, hx_members) BIND(C, name="letkf_gpu_set_obs_members")
       IMPORT :: c_int, c_ptr, c_float, c_int32_t, c_int64_t
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int64_t), VALUE :: n_obs
       REAL(c_float), INTENT(in) :: lat(*), lon(*), zdim(*), time(*), val(*), err(*)
       INTEGER(c_int32_t), INTENT(in) :: obsid(*), platid(*), qc(*)
       REAL(c_float), INTENT(in) :: hx_members(*)
       INTEGER(c_int) :: letkf_gpu_set_obs_members
     END FUNCTION letkf_gpu_set_obs_members

     FUNCTION letkf_gpu_get_state_members(h, members, mean_ij, sprd_ij) &
          BIND(C, name="letkf_gpu_get_state_members")
       IMPORT :: c_int, c_ptr, c_float
       TYPE(c_ptr), VALUE :: h
       REAL(c_float), INTENT(out) :: members(*), mean_ij(*), sprd_ij(*)
       INTEGER(c_int) :: letkf_gpu_get_state_members
     END FUNCTION letkf_gpu_get_state_members

     FUNCTION letkf_gpu_autotune(h, repeats, t) BIND(C, name="letkf_gpu_autotune")
       IMPORT :: c_int, c_ptr, c_int32_t, letkf_gpu_tuning
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int32_t), VALUE :: repeats
       TYPE(letkf_gpu_tuning), INTENT(out) :: t
       INTEGER(c_int) :: letkf_gpu_autotune
     END FUNCTION letkf_gpu_autotune

     FUNCTION letkf_gpu_get_tuning(h, t) BIND(C, name="letkf_gpu_get_tuning")
       IMPORT :: c_int, c_ptr, letkf_gpu_tuning
       TYPE(c_ptr), VALUE :: h
       TYPE(letkf_gpu_tuning), INTENT(out) :: t
       INTEGER(c_int) :: letkf_gpu_get_tuning
     END FUNCTION letkf_gpu_get_tuning

     FUNCTION letkf_gpu_set_tuning(h, t) BIND(C, name="letkf_gpu_set_tuning")
       IMPORT :: c_int, c_ptr, letkf_gpu_tuning
       TYPE(c_ptr), VALUE :: h
       TYPE(letkf_gpu_tuning), INTENT(in) :: t
       INTEGER(c_int) :: letkf_gpu_set_tuning
     END FUNCTION letkf_gpu_set_tuning

     FUNCTION c_strlen(s) BIND(C, name="strlen")
       IMPORT :: c_ptr, c_size_t
       TYPE(c_ptr), VALUE :: s
       INTEGER(c_size_t) :: c_strlen
     END FUNCTION c_strlen
  END INTERFACE

CONTAINS

  !> message of the last failure on h (or of the last failed create when h
  !! is c_null_ptr) as a Fortran string
  FUNCTION letkf_gpu_errmsg(h) RESULT(msg)
    TYPE(c_ptr), INTENT(in) :: h
    CHARACTER(len=:), ALLOCATABLE :: msg
    TYPE(c_ptr) :: p
    CHARACTER(kind=c_char), POINTER :: chars(:)
    INTEGER :: n, i
    p = letkf_gpu_last_error(h)
    IF (.NOT. c_associated(p)) THEN
       msg = ""
       RETURN
    END IF
    n = INT(c_strlen(p))
    CALL c_f_pointer(p, chars, [n])
    ALLOCATE(CHARACTER(len=n) :: msg)
    DO i = 1, n
       msg(i:i) = chars(i)
    END DO
  END FUNCTION letkf_gpu_errmsg

END MODULE letkf_gpu_mod
