!> @file letkf_gpu_loc.F90
!! Localizer settings of the reference turned into the GPU library's
!! descriptor.  Self-contained (no module of the reference is used), so the
!! localizer plugins, the solver host and the stand-alone driver all call
!! the same code.
!!
!! to_c_table      a linearinterp_lat table (the npoints, lat(:) and dist(:)
!!                 of src/letkf_loc.F90:126-133) into letkf_linearinterp_lat;
!!                 a completed table (0 and 90 degree end points already
!!                 added by linearinterp_lat_init, :275-293) may be passed as
!!                 it is, since adding the end points again changes nothing
!! vgauss_vweight  the vertical factor of LOC_VGAUSS for one observation and
!!                 one level, exp(-(zdim - depth)^2 / (2 scale^2)), in double
!!                 precision and rounded to REAL like the horizontal weight
!! vgauss_desc     the letkf_loc_desc of a LOC_VGAUSS localizer: horizontal
!!                 table, vertical scale and the level depths (one
!!                 localization group per level); the descriptor points at
!!                 depth, which must stay in place until letkf_gpu_set_localizer
!!                 has copied it
MODULE letkf_gpu_loc
  USE, INTRINSIC :: iso_c_binding
  USE letkf_gpu_mod
  IMPLICIT NONE
  PRIVATE

  PUBLIC :: to_c_table, vgauss_vweight, vgauss_desc

CONTAINS

  SUBROUTINE to_c_table(lat, dist, tab, ierr)
    REAL(c_float), INTENT(in) :: lat(:), dist(:)
    TYPE(letkf_linearinterp_lat), INTENT(out) :: tab
    INTEGER, INTENT(out) :: ierr
    INTEGER :: n
    n = SIZE(lat)
    ierr = LETKF_OK
    tab%npoints = 0
    tab%lat = 0.0
    tab%radius = 0.0
    IF (n /= SIZE(dist) .OR. n < 1 .OR. n > LETKF_MAX_INTERP) THEN
       ierr = LETKF_EINVAL
       RETURN
    END IF
    tab%npoints = n
    tab%lat(1:n) = lat
    tab%radius(1:n) = dist
  END SUBROUTINE to_c_table


  PURE FUNCTION vgauss_vweight(zdim, depth, scale) RESULT(w)
    REAL(c_float), INTENT(in) :: zdim, depth, scale
    REAL(c_float) :: w
    REAL(c_double) :: dz, s
    dz = REAL(zdim, c_double) - REAL(depth, c_double)
    s = REAL(scale, c_double)
    w = REAL(EXP(-(dz * dz) / (2.0_c_double * s * s)), c_float)
  END FUNCTION vgauss_vweight


  SUBROUTINE vgauss_desc(hz_lat, hz_dist, scale, depth, desc, ierr)
    REAL(c_float), INTENT(in) :: hz_lat(:), hz_dist(:), scale
    REAL(c_float), INTENT(in), TARGET :: depth(:)
    TYPE(letkf_loc_desc), INTENT(out) :: desc
    INTEGER, INTENT(out) :: ierr
    desc%cls = LETKF_LOC_VGAUSS
    CALL to_c_table(hz_lat, hz_dist, desc%hzloc, ierr)
    IF (ierr /= LETKF_OK) RETURN
    IF (.NOT. (scale > 0.0) .OR. SIZE(depth) < 1) THEN
       ierr = LETKF_EINVAL
       RETURN
    END IF
    desc%vt_scale = scale
    desc%nlev = INT(SIZE(depth), c_int32_t)
    desc%lev_depth = c_loc(depth(1))
  END SUBROUTINE vgauss_desc

END MODULE letkf_gpu_loc
