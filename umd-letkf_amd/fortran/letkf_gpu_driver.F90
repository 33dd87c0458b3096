!> @file letkf_gpu_driver.F90
!! Stand-alone Fortran host: reads an analysis problem from a binary file,
!! runs it through letkf_solver_gpu (init / run / final) and writes the
!! analysis and the selection diagnostics.
!!
!!   letkf_gpu_driver --sizes            C-interop sizes of the ABI structs
!!   letkf_gpu_driver --vweight z d s    vgauss_vweight(z, d, s) of letkf_gpu_loc
!!   letkf_gpu_driver <problem> <result>
!!
!! LOC_NOVRT and LOC_VGAUSS descriptors are rebuilt from the file's table,
!! scale and depths through letkf_gpu_loc (to_c_table, vgauss_desc), as a
!! localizer plugin of the reference builds them.
!!
!! Problem file (stream, little endian):
!!   int64  k, ns, nij, nobs
!!   int32  precision, path, nvars, nlevdepth
!!   real32 rtps, rtpp, mul
!!   bytes  letkf_loc_desc image (its lev_depth pointer is replaced)
!!   real32 lev_depth(nlevdepth)
!!   real32 state(k,ns,nij), mean(ns,nij), sprd(ns,nij), lat(nij), lon(nij)
!!   int32  mask(nij)
!!   int32  grid_s_idx(nvars), levels(nvars)
!!   real32 ana_inc_max(nvars), ana_bounds(2,nvars)
!!   real32 olat, olon, ozdim, otime, oval, oerr (nobs each)
!!   int32  obsid(nobs), platid(nobs)
!!   real32 hx(k,nobs), hx_mean(nobs)
!! Result file:
!!   int32  maxgroups
!!   real32 state(k,ns,nij), mean(ns,nij), sprd(ns,nij)   (in place, FP32)
!!   real64 state, mean, sprd                              (letkf_gpu_get_state_f64)
!!   real32 lg_obscount(nij,maxgroups)
!!   int64  lg_selhash(nij,maxgroups)
PROGRAM letkf_gpu_driver
  USE, INTRINSIC :: iso_c_binding
  USE letkf_gpu_mod
  USE letkf_gpu_loc
  USE letkf_solver_gpu
  IMPLICIT NONE

  CHARACTER(len=1024) :: fin, fout
  INTEGER(c_int64_t) :: k, ns, nij, nobs
  INTEGER(c_int32_t) :: prec, path, nvars, nlevd, maxg
  REAL(c_float) :: rtps, rtpp, mul
  TYPE(letkf_loc_desc) :: loc, loc2
  CHARACTER(len=64) :: a1, a2, a3
  REAL(c_float) :: z1, z2, z3
  TYPE(letkf_gpu_dims) :: dims
  TYPE(letkf_linearinterp_lat) :: tab
  TYPE(letkf_gpu_diag) :: diag
  TYPE(letkf_gpu_timing) :: timing
  TYPE(letkf_gpu_tuning) :: tuning
  INTEGER(c_int8_t), ALLOCATABLE :: raw(:)
  REAL(c_float), ALLOCATABLE, TARGET :: lev_depth(:)
  REAL(c_float), ALLOCATABLE :: state(:,:,:), mean(:,:), sprd(:,:), lat(:), lon(:)
  INTEGER(c_int32_t), ALLOCATABLE :: imask(:), s_idx(:), levels(:), obsid(:), platid(:)
  LOGICAL, ALLOCATABLE :: mask(:)
  REAL(c_float), ALLOCATABLE :: inc_max(:), bounds(:,:)
  REAL(c_float), ALLOCATABLE :: olat(:), olon(:), ozdim(:), otime(:), oval(:), oerr(:)
  REAL(c_float), ALLOCATABLE :: hx(:,:), hxm(:)
  REAL(c_double), ALLOCATABLE :: st64(:,:,:), mn64(:,:), sp64(:,:)
  REAL(c_float), ALLOCATABLE, TARGET :: lgm(:,:)
  INTEGER(c_int64_t), ALLOCATABLE, TARGET :: lgh(:,:)
  INTEGER :: u, ierr

  CALL get_command_argument(1, fin)
  IF (TRIM(fin) == "--sizes") THEN
     PRINT '(A,6I8)', "sizes", c_sizeof(dims), c_sizeof(tab), c_sizeof(loc), &
          c_sizeof(diag), c_sizeof(timing), c_sizeof(tuning)
     STOP
  END IF
  IF (TRIM(fin) == "--vweight") THEN
     CALL get_command_argument(2, a1)
     CALL get_command_argument(3, a2)
     CALL get_command_argument(4, a3)
     READ(a1, *) z1
     READ(a2, *) z2
     READ(a3, *) z3
     PRINT '(A,ES24.16)', "vweight", REAL(vgauss_vweight(z1, z2, z3), c_double)
     STOP
  END IF
  CALL get_command_argument(2, fout)

  OPEN(newunit=u, file=TRIM(fin), access="stream", form="unformatted", status="old", &
       action="read")
  READ(u) k, ns, nij, nobs
  READ(u) prec, path, nvars, nlevd
  READ(u) rtps, rtpp, mul
  ALLOCATE(raw(c_sizeof(loc)))
  READ(u) raw
  loc = TRANSFER(raw, loc)
  ALLOCATE(lev_depth(MAX(nlevd, 1)))
  IF (nlevd > 0) READ(u) lev_depth(1:nlevd)
  loc%lev_depth = c_null_ptr
  IF (nlevd > 0) loc%lev_depth = c_loc(lev_depth)
  IF (loc%cls == LETKF_LOC_VGAUSS) THEN
     CALL vgauss_desc(loc%hzloc%lat(1:loc%hzloc%npoints), loc%hzloc%radius(1:loc%hzloc%npoints), &
          loc%vt_scale, lev_depth(1:nlevd), loc2, ierr)
     CALL check("vgauss_desc")
     loc = loc2
  ELSE IF (loc%cls == LETKF_LOC_NOVRT) THEN
     loc2%cls = LETKF_LOC_NOVRT
     CALL to_c_table(loc%hzloc%lat(1:loc%hzloc%npoints), loc%hzloc%radius(1:loc%hzloc%npoints), &
          loc2%hzloc, ierr)
     CALL check("to_c_table")
     loc = loc2
  END IF
  ALLOCATE(state(k, ns, nij), mean(ns, nij), sprd(ns, nij), lat(nij), lon(nij), imask(nij))
  READ(u) state, mean, sprd, lat, lon, imask
  ALLOCATE(mask(nij))
  mask = imask /= 0
  ALLOCATE(s_idx(nvars), levels(nvars), inc_max(nvars), bounds(2, nvars))
  READ(u) s_idx, levels, inc_max, bounds
  ALLOCATE(olat(nobs), olon(nobs), ozdim(nobs), otime(nobs), oval(nobs), oerr(nobs))
  ALLOCATE(obsid(nobs), platid(nobs), hx(k, nobs), hxm(nobs))
  READ(u) olat, olon, ozdim, otime, oval, oerr, obsid, platid, hx, hxm
  CLOSE(u)

  CALL letkf_solver_gpu_init(INT(k), INT(ns), INT(nij), prec, path, loc, s_idx, levels, &
       inc_max, bounds, rtps, rtpp, mul, ierr)
  CALL check("letkf_solver_gpu_init")
  CALL letkf_solver_gpu_run(state, mean, sprd, lat, lon, mask, olat, olon, ozdim, otime, &
       oval, oerr, obsid, platid, hx, hxm, ierr)
  CALL check("letkf_solver_gpu_run")

  ierr = letkf_gpu_maxgroups(letkf_solver_gpu_handle(), maxg)
  CALL check("letkf_gpu_maxgroups")
  ALLOCATE(st64(k, ns, nij), mn64(ns, nij), sp64(ns, nij), lgm(nij, maxg), lgh(nij, maxg))
  ierr = letkf_gpu_get_state_f64(letkf_solver_gpu_handle(), st64, mn64, sp64)
  CALL check("letkf_gpu_get_state_f64")
  diag%lg_obscount = c_loc(lgm)
  diag%lg_selhash = c_loc(lgh)
  ierr = letkf_gpu_get_diag(letkf_solver_gpu_handle(), diag)
  CALL check("letkf_gpu_get_diag")
  CALL letkf_solver_gpu_final(ierr)
  CALL check("letkf_solver_gpu_final")

  OPEN(newunit=u, file=TRIM(fout), access="stream", form="unformatted", status="replace", &
       action="write")
  WRITE(u) maxg, state, mean, sprd, st64, mn64, sp64, lgm, lgh
  CLOSE(u)
  PRINT '(A)', "letkf_gpu_driver: done"

CONTAINS

  SUBROUTINE check(what)
    CHARACTER(len=*), INTENT(in) :: what
    IF (ierr /= LETKF_OK) THEN
       PRINT '(A,A,I0,A,A)', what, ": status ", ierr, ": ", letkf_gpu_errmsg(letkf_solver_gpu_handle())
       STOP 1
    END IF
  END SUBROUTINE check

END PROGRAM letkf_gpu_driver
