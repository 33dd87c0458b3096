!> @file letkf_loc_vgauss.F90
!! LOC_VGAUSS: a localizer plugin for the reference's own build, so that the
!! vertical-Gaussian localization the GPU library offers can also be chosen
!! from the unchanged YAML and run on the CPU loop:
!!
!!   localization:
!!     class: loc_vgauss
!!     hzloc: {type: linearinterp_lat, value: [{lat: 0.0, radius: 700.0e3}, ...]}
!!     vtloc: 300.0        # Gaussian scale of the vertical weight, zdim units
!!
!! Weight of an observation for level group l: the horizontal Gaspari-Cohn
!! of LOC_NOVRT (letkf_loc_gc with the hzloc radius at the gridpoint's
!! latitude) times exp(-(zdim - depth_l)^2 / (2 vtloc^2)), depth_l the
!! nominal depth of level l of the first vertical grid.  Group l holds the
!! level-l slab of every state variable that has at least l levels.
!!
!! It implements the deferred procedures of letkf_localizer
!! (src/letkf_loc.F90:54-110) like letkf_loc_novrt does, and is registered
!! through add_letkf_plugins (src/letkf_pluginloader.cmake:21): add it, with
!! letkf_gpu_mod.F90 and letkf_gpu_loc.F90, to the LETKF_PLUGINS list of
!! src/subclass.letkf_localizer/CMakeLists.txt.  It uses the reference's
!! modules, so it is compiled only inside the reference's tree.
MODULE letkf_loc_vgauss_mod
  USE letkf_config
  USE letkf_mpi
  USE letkf_loc
  USE letkf_obs
  USE letkf_state
  USE letkf_gpu_loc, ONLY: vgauss_vweight
  IMPLICIT NONE
  PRIVATE

  TYPE, EXTENDS(letkf_localizer), PUBLIC :: letkf_loc_vgauss
     TYPE(linearinterp_lat) :: hzdist
     REAL :: vtloc = 0.0
     REAL, ALLOCATABLE :: depth(:)
   CONTAINS
     PROCEDURE, NOPASS :: name => loc_vgauss_name
     PROCEDURE, NOPASS :: desc => loc_vgauss_desc
     PROCEDURE         :: init => loc_vgauss_init
     PROCEDURE         :: final => loc_vgauss_final
     PROCEDURE         :: groups => loc_vgauss_groups
     PROCEDURE         :: localize => loc_vgauss_localize
     PROCEDURE         :: maxhz => loc_vgauss_maxhz
  END TYPE letkf_loc_vgauss

CONTAINS

  FUNCTION loc_vgauss_name() RESULT(str)
    CHARACTER(:), ALLOCATABLE :: str
    str = "LOC_VGAUSS"
  END FUNCTION loc_vgauss_name


  FUNCTION loc_vgauss_desc() RESULT(str)
    CHARACTER(:), ALLOCATABLE :: str
    str = "Horizontal Gaspari-Cohn times a vertical Gaussian, one group per level"
  END FUNCTION loc_vgauss_desc


  SUBROUTINE loc_vgauss_init(self, config)
    CLASS(letkf_loc_vgauss), INTENT(inout) :: self
    TYPE(configuration), INTENT(in) :: config
    CHARACTER(:), ALLOCATABLE :: str
    TYPE(configuration) :: config2, config3

    CALL config%get("hzloc", config2)
    CALL config2%get("type", str)
    IF (str == "linearinterp_lat") THEN
       CALL config2%get("value", config3)
       self%hzdist = linearinterp_lat(config3)
    ELSE
       CALL letkf_mpi_abort('Unrecognized hzloc type: "'//TRIM(str)//'"')
    END IF
    CALL config%get("vtloc", self%vtloc)
    IF (self%vtloc <= 0.0) CALL letkf_mpi_abort("LOC_VGAUSS: vtloc must be > 0")
    IF (SIZE(vtgrids) < 1) CALL letkf_mpi_abort("LOC_VGAUSS: no vertical grid")
    ALLOCATE(self%depth(SIZE(vtgrids(1)%vert_nom)))
    self%depth = vtgrids(1)%vert_nom
    IF (MAXVAL(statevars(:)%levels) > SIZE(self%depth)) &
         CALL letkf_mpi_abort("LOC_VGAUSS: a variable has more levels than the vertical grid")
    self%maxgroups = SIZE(self%depth)
    IF (pe_isroot) THEN
       PRINT *, ""
       PRINT *, "LOC_VGAUSS initialization"
       PRINT *, "------------------------------------------------------------"
       PRINT *, "hzloc="
       CALL self%hzdist%print()
       PRINT *, "vtloc=", self%vtloc, " levels=", SIZE(self%depth)
    END IF
  END SUBROUTINE loc_vgauss_init


  SUBROUTINE loc_vgauss_final(self)
    CLASS(letkf_loc_vgauss), INTENT(inout) :: self
    IF (ALLOCATED(self%depth)) DEALLOCATE(self%depth)
  END SUBROUTINE loc_vgauss_final


  SUBROUTINE loc_vgauss_groups(self, ij, groups)
    CLASS(letkf_loc_vgauss), INTENT(inout) :: self
    INTEGER, INTENT(in)  :: ij
    TYPE(letkf_localizer_group), ALLOCATABLE, INTENT(inout) :: groups(:)
    INTEGER :: l, v, n

    ALLOCATE(groups(SIZE(self%depth)))
    DO l = 1, SIZE(self%depth)
       groups(l)%id = l
       n = COUNT(statevars(:)%levels >= l)
       ALLOCATE(groups(l)%slab(n))
       n = 0
       DO v = 1, SIZE(statevars)
          IF (statevars(v)%levels >= l) THEN
             n = n + 1
             groups(l)%slab(n) = statevars(v)%grid_s_idx + l - 1
          END IF
       END DO
    END DO
  END SUBROUTINE loc_vgauss_groups


  FUNCTION loc_vgauss_maxhz(self, ij) RESULT(dist)
    CLASS(letkf_loc_vgauss), INTENT(inout) :: self
    INTEGER, INTENT(in)  :: ij
    REAL :: dist
    dist = self%hzdist%get_dist(lat_ij(ij))
  END FUNCTION loc_vgauss_maxhz


  FUNCTION loc_vgauss_localize(self, ij, group, obs, dist) RESULT(loc)
    CLASS(letkf_loc_vgauss), INTENT(inout) :: self
    INTEGER, INTENT(in) :: ij
    TYPE(letkf_localizer_group), INTENT(in) :: group
    TYPE(letkf_observation), INTENT(in) :: obs
    REAL, INTENT(in) :: dist
    REAL :: loc
    loc = letkf_loc_gc(dist, self%hzdist%get_dist(lat_ij(ij)))
    loc = loc * vgauss_vweight(obs%zdim, self%depth(group%id), self%vtloc)
  END FUNCTION loc_vgauss_localize

END MODULE letkf_loc_vgauss_mod
