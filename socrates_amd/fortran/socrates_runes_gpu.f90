! Fortran binding of the MI355X two-stream flux library (libsocgpu.so).
!
! Provides the module names, procedure names and keyword arguments of the
! reference interface for the two-stream flux path, so that a host model that
! calls SOCRATES through `runes` compiles against this file instead of the
! reference's interface_core:
!   realtype_rd            kinds RealK, RealExt          (realtype_rd.f90:14-20)
!   def_spectrum           StrSpecData (a handle here)   (def_spectrum.F90:33)
!   socrates_def_diag      StrDiag, the path's fields    (socrates_def_diag.F90:31-175)
!   socrates_set_spectrum  set_spectrum, set_mcica       (socrates_set_spectrum.F90:29, 551)
!   socrates_runes         runes, StrDiag, ip_* codes    (socrates_runes.F90:47-683)
! runes takes both of the reference's calling forms: full fields
! (n_profile, n_layer), and the host-model form of strided one-dimensional
! fields with a profile list (profile_list, n_layer_stride, n_level_stride,
! *_1d, l_profile_last, l_invert), every gas as a field, a strided field or a
! well-mixed ratio with the l_*_well_mixed switches. The binding gathers the
! caller's arrays into the library's layout exactly as set_atm, set_bound,
! set_cld and set_cld_mcica index them (socrates_set_atm.F90:191-622,
! socrates_set_bound.F90:112-290, socrates_set_cld.F90:861-910,
! socrates_set_cld_mcica.F90:159-167) and scatters the outputs as set_diag
! does (socrates_set_diag.F90:104-256, 549-623).
! Modal aerosols (n_aer_mode, aer_* and their _1d forms, l_aerosol_mode)
! are taken with every cloud representation. Arguments of runes outside
! the path (CLASSIC aerosols, tiles, orography, cloud sub-fractions,
! photolysis, the clean-air diagnostics, ...) are not
! declared: a call naming one does not compile, rather than being silently
! ignored. Ice takes the spectral file's ice types of Fu's phase-function
! fit, Baran's fit and Pade-2; without i_st_ice a cloudy call has no ice
! component (the reference requires one). A gas the spectral file lists is
! used as the reference uses it; one the library cannot take (other than
! H2O, CO2, O3, N2O, CH4, O2, SO2, OCS, CFC-11, CFC-12, CFC-113, HCFC-22 and
! HFC-134a) stops the call if it is given a non-zero amount.
!
! Errors: the library returns a status; a non-zero status is reported here
! and stops the program, which is the reference's ereport(i_err_fatal)
! behaviour (ereport_mod.F90:27-40).
module realtype_rd
  implicit none
  integer, parameter :: RealK = selected_real_kind(15, 307)
  integer, parameter :: RealExt = selected_real_kind(15, 307)
end module realtype_rd


module socrates_gpu_c
  use, intrinsic :: iso_c_binding
  implicit none

  ! SocgpuRunesArgs of include/socgpu.h, field for field
  type, bind(C) :: socgpu_runes_args
    integer(c_int) :: n_profile = 0, n_layer = 0, i_source = 0, l_rayleigh = 0
    integer(c_int) :: i_cloud_representation = 0, i_overlap = 0, i_inhom = 0
    integer(c_int) :: i_mcica_sampling = 0, i_st_water = 5, i_cnv_water = 5
    integer(c_int) :: n_subcol_gen = 0, l_inputs_on_device = 0
    real(c_double) :: co2_mix_ratio = 0.0_c_double, cloud_vertical_decorr = 0.0_c_double
    real(c_double) :: conv_vertical_decorr = 0.0_c_double, cloud_horizontal_rsd = 0.0_c_double
    integer(c_int) :: l_grey_albedo = 0
    real(c_double) :: grey_albedo = 0.0_c_double
    type(c_ptr) :: p_layer = c_null_ptr, t_layer = c_null_ptr, mass = c_null_ptr
    type(c_ptr) :: density = c_null_ptr, layer_heat_capacity = c_null_ptr
    type(c_ptr) :: h2o = c_null_ptr, o3 = c_null_ptr
    type(c_ptr) :: p_level = c_null_ptr, t_level = c_null_ptr
    type(c_ptr) :: t_ground = c_null_ptr, cos_zenith_angle = c_null_ptr
    type(c_ptr) :: solar_irrad = c_null_ptr
    type(c_ptr) :: albedo_diff = c_null_ptr, albedo_dir = c_null_ptr
    type(c_ptr) :: cloud_frac = c_null_ptr, liq_mmr = c_null_ptr, liq_dim = c_null_ptr
    type(c_ptr) :: conv_frac = c_null_ptr, liq_conv_mmr = c_null_ptr
    type(c_ptr) :: liq_conv_dim = c_null_ptr
    type(c_ptr) :: rand_seed = c_null_ptr
    ! n2o, ch4, o2, so2, ocs, cfc11, cfc12, cfc113, hcfc22, hfc134a
    real(c_double) :: gas_mix_ratio(10) = 0.0_c_double
    integer(c_int) :: l_flux_ground = 0, l_co2_well_mixed = 0
    type(c_ptr) :: flux_ground = c_null_ptr, co2 = c_null_ptr
    real(c_double) :: h2o_mix_ratio = 0.0_c_double, o3_mix_ratio = 0.0_c_double
    ! fields of the further gases, in gas_mix_ratio's order
    type(c_ptr) :: gas_field(10) = c_null_ptr
    integer(c_int) :: i_st_ice = 0, i_cnv_ice = 0
    type(c_ptr) :: ice_mmr = c_null_ptr, ice_dim = c_null_ptr
    type(c_ptr) :: ice_conv_mmr = c_null_ptr, ice_conv_dim = c_null_ptr
    integer(c_int) :: l_ir_source_linear = 0, l_no_continuum = 0
    integer(c_int) :: i_scatter_method = 0
    ! modal aerosols: mixing ratio (profile, layer, mode), optical properties
    ! (profile, layer, mode, band)
    integer(c_int) :: n_aer_mode = 0
    type(c_ptr) :: aer_mix_ratio = c_null_ptr, aer_absorption = c_null_ptr
    type(c_ptr) :: aer_scattering = c_null_ptr, aer_asymmetry = c_null_ptr
  end type socgpu_runes_args

  ! SocgpuDiag
  type, bind(C) :: socgpu_diag
    type(c_ptr) :: flux_up = c_null_ptr, flux_down = c_null_ptr
    type(c_ptr) :: flux_direct = c_null_ptr, heating_rate = c_null_ptr
    integer(c_int) :: l_outputs_on_device = 0
    type(c_ptr) :: flux_up_clear = c_null_ptr, flux_down_clear = c_null_ptr
    type(c_ptr) :: flux_direct_clear = c_null_ptr
  end type socgpu_diag

  interface
    integer(c_int) function socgpu_init(device) bind(C, name='socgpu_init')
      import :: c_int
      integer(c_int), value :: device
    end function socgpu_init
    integer(c_int) function socgpu_spectrum_register_gases(name, path, n_type, &
        type_absorb) bind(C, name='socgpu_spectrum_register_gases')
      import :: c_char, c_int
      character(kind=c_char), dimension(*) :: name, path
      integer(c_int), value :: n_type
      integer(c_int) :: type_absorb(*)
    end function socgpu_spectrum_register_gases
    integer(c_int) function socgpu_spectrum_query(name, n_band, n_absorb, &
        type_absorb, max_absorb) bind(C, name='socgpu_spectrum_query')
      import :: c_char, c_int
      character(kind=c_char), dimension(*) :: name
      integer(c_int) :: n_band, n_absorb
      integer(c_int) :: type_absorb(*)
      integer(c_int), value :: max_absorb
    end function socgpu_spectrum_query
    integer(c_int) function socgpu_mcica_register(path, sw_name, lw_name) &
        bind(C, name='socgpu_mcica_register')
      import :: c_char, c_int, c_ptr
      character(kind=c_char), dimension(*) :: path
      type(c_ptr), value :: sw_name, lw_name
    end function socgpu_mcica_register
    integer(c_int) function socgpu_runes(name, args, diag, precision) &
        bind(C, name='socgpu_runes')
      import :: c_char, c_int, socgpu_runes_args, socgpu_diag
      character(kind=c_char), dimension(*) :: name
      type(socgpu_runes_args) :: args
      type(socgpu_diag) :: diag
      integer(c_int), value :: precision
    end function socgpu_runes
    type(c_ptr) function socgpu_last_error() bind(C, name='socgpu_last_error')
      import :: c_ptr
    end function socgpu_last_error
  end interface

  logical, save :: initialised = .false.

contains

  function c_str(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s) + 1) :: c
    c = trim(s) // c_null_char
  end function c_str

  ! Message of the last library failure on this thread
  function last_error() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: i
    msg = ' '
    cp = socgpu_last_error()
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [512])
    do i = 1, 512
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function last_error

  ! Stop on a non-zero status, as ereport(i_err_fatal) does in the reference
  subroutine check(status, routine)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: routine
    if (status /= 0) then
      write(*, '(a, a, a, i0, a, a)') '? Error from routine: ', routine, &
        ' status ', status, ': ', trim(last_error())
      error stop 1
    end if
  end subroutine check

  subroutine ensure_init()
    if (.not. initialised) then
      call check(socgpu_init(0_c_int), 'SOCGPU_INIT')
      initialised = .true.
    end if
  end subroutine ensure_init

end module socrates_gpu_c


! The spectrum handle: a spectral file registered with the library under a
! name of its own. set_spectrum(spectrum=...) fills it; runes(spectrum=...)
! uses it. The reference's StrSpecData holds the file's tables
! (def_spectrum.F90:33-472); here the tables stay on the device.
module def_spectrum
  implicit none
  private
  public :: StrSpecData
  type :: StrSpecData
    character(len=64) :: name = ''
    character(len=1024) :: file = ''
    integer :: n_keep = -1          ! absorbers kept (-1: all) and their identifiers
    integer :: keep(80) = 0
  end type StrSpecData
end module def_spectrum


module socrates_set_spectrum
  use, intrinsic :: iso_c_binding
  use socrates_gpu_c
  use def_spectrum, only: StrSpecData
  implicit none
  private
  public :: set_spectrum, set_mcica
  integer, save :: n_anonymous = 0

contains

  ! Reads a spectral file and keeps it on the device under spectrum_name
  ! and/or in the handle spectrum (socrates_set_spectrum.F90:29-149): with
  ! both, the named slot takes the handle's spectrum. Without l_all_gases,
  ! only the absorbers flagged l_<gas> = .true. are kept (compress_spectrum,
  ! :152-321). n_instances is accepted for compatibility (the library has no
  ! slot limit); wavelength_blue only weights the blue-band diagnostics,
  ! which are not on this path.
  subroutine set_spectrum(n_instances, spectrum, spectrum_name, spectral_file, &
      l_h2o, l_co2, l_o3, l_n2o, l_co, l_ch4, l_o2, l_no, l_so2, l_no2, &
      l_nh3, l_hno3, l_n2, l_cfc11, l_cfc12, l_cfc113, l_hcfc22, l_hfc125, &
      l_hfc134a, l_cfc114, l_tio, l_vo, l_h2, l_he, l_ocs, l_na, l_k, &
      l_feh, l_crh, l_li, l_rb, l_cs, l_ph3, l_c2h2, l_hcn, l_h2s, l_ar, &
      l_o, l_n, l_no3, l_n2o5, l_hono, l_ho2no2, l_h2o2, l_c2h6, l_ch3, &
      l_h2co, l_ho2, l_hdo, l_hcl, l_hf, l_cosso, l_tosso, l_yosos, &
      l_ch3cho, l_ch3ooh, l_ch3coch3, l_ch3cocho, l_chocho, l_c2h5cho, &
      l_hoch2cho, l_c2h5coch3, l_mvk, l_macr, l_pan, l_ch3ono2, &
      l_all_gases, wavelength_blue)
    use realtype_rd, only: RealExt
    integer, intent(in), optional :: n_instances
    type(StrSpecData), intent(inout), target, optional :: spectrum
    character(len=*), intent(in), optional :: spectrum_name
    character(len=*), intent(in), optional :: spectral_file
    logical, intent(in), optional :: &
      l_h2o, l_co2, l_o3, l_n2o, l_co, l_ch4, l_o2, l_no, l_so2, l_no2, &
      l_nh3, l_hno3, l_n2, l_cfc11, l_cfc12, l_cfc113, l_hcfc22, l_hfc125, &
      l_hfc134a, l_cfc114, l_tio, l_vo, l_h2, l_he, l_ocs, l_na, l_k, &
      l_feh, l_crh, l_li, l_rb, l_cs, l_ph3, l_c2h2, l_hcn, l_h2s, l_ar, &
      l_o, l_n, l_no3, l_n2o5, l_hono, l_ho2no2, l_h2o2, l_c2h6, l_ch3, &
      l_h2co, l_ho2, l_hdo, l_hcl, l_hf, l_cosso, l_tosso, l_yosos, &
      l_ch3cho, l_ch3ooh, l_ch3coch3, l_ch3cocho, l_chocho, l_c2h5cho, &
      l_hoch2cho, l_c2h5coch3, l_mvk, l_macr, l_pan, l_ch3ono2
    logical, intent(in), optional :: l_all_gases
    real(RealExt), intent(in), optional :: wavelength_blue
    integer(c_int) :: keep(80)
    integer :: n_keep
    logical :: all_gases
    character(len=64) :: name

    if (.not. (present(spectrum_name) .or. present(spectrum))) then
      if (present(n_instances)) return
      call check(-1_c_int, 'SET_SPECTRUM')
    end if
    call ensure_init()
    if (present(spectrum_name) .and. present(spectrum)) then
      ! the named slot takes the handle's spectrum
      if (spectrum%name == '') call check(-1_c_int, 'SET_SPECTRUM')
      keep = int(spectrum%keep, c_int)
      call check(socgpu_spectrum_register_gases(c_str(spectrum_name), c_str(spectrum%file), &
                 int(spectrum%n_keep, c_int), keep), 'SET_SPECTRUM')
      return
    end if
    if (.not. present(spectral_file)) call check(-1_c_int, 'SET_SPECTRUM')
    all_gases = .false.
    if (present(l_all_gases)) all_gases = l_all_gases
    n_keep = 0
    keep = 0
    if (all_gases) then
      n_keep = -1
    else
      call flag(l_h2o, 1)
      call flag(l_co2, 2)
      call flag(l_o3, 3)
      call flag(l_n2o, 4)
      call flag(l_co, 5)
      call flag(l_ch4, 6)
      call flag(l_o2, 7)
      call flag(l_no, 8)
      call flag(l_so2, 9)
      call flag(l_no2, 10)
      call flag(l_nh3, 11)
      call flag(l_hno3, 12)
      call flag(l_n2, 13)
      call flag(l_cfc11, 14)
      call flag(l_cfc12, 15)
      call flag(l_cfc113, 16)
      call flag(l_hcfc22, 17)
      call flag(l_hfc125, 18)
      call flag(l_hfc134a, 19)
      call flag(l_cfc114, 20)
      call flag(l_tio, 21)
      call flag(l_vo, 22)
      call flag(l_h2, 23)
      call flag(l_he, 24)
      call flag(l_ocs, 25)
      call flag(l_na, 26)
      call flag(l_k, 27)
      call flag(l_feh, 28)
      call flag(l_crh, 29)
      call flag(l_li, 30)
      call flag(l_rb, 31)
      call flag(l_cs, 32)
      call flag(l_ph3, 33)
      call flag(l_c2h2, 34)
      call flag(l_hcn, 35)
      call flag(l_h2s, 36)
      call flag(l_ar, 37)
      call flag(l_o, 39)
      call flag(l_n, 40)
      call flag(l_no3, 41)
      call flag(l_n2o5, 42)
      call flag(l_hono, 43)
      call flag(l_ho2no2, 44)
      call flag(l_h2o2, 45)
      call flag(l_c2h6, 46)
      call flag(l_ch3, 47)
      call flag(l_h2co, 48)
      call flag(l_ho2, 49)
      call flag(l_hdo, 50)
      call flag(l_hcl, 51)
      call flag(l_hf, 52)
      call flag(l_cosso, 53)
      call flag(l_tosso, 54)
      call flag(l_yosos, 55)
      call flag(l_ch3cho, 56)
      call flag(l_ch3ooh, 57)
      call flag(l_ch3coch3, 58)
      call flag(l_ch3cocho, 59)
      call flag(l_chocho, 60)
      call flag(l_c2h5cho, 61)
      call flag(l_hoch2cho, 62)
      call flag(l_c2h5coch3, 63)
      call flag(l_mvk, 64)
      call flag(l_macr, 65)
      call flag(l_pan, 66)
      call flag(l_ch3ono2, 67)
    end if
    if (present(spectrum_name)) then
      name = spectrum_name
    else
      n_anonymous = n_anonymous + 1
      write(name, '(a, i0)') '.spectrum.', n_anonymous
      spectrum%name = name
      spectrum%file = spectral_file
      spectrum%n_keep = n_keep
      spectrum%keep = keep
    end if
    call check(socgpu_spectrum_register_gases(c_str(name), c_str(spectral_file), &
               int(n_keep, c_int), keep), 'SET_SPECTRUM')

  contains

    subroutine flag(l, ip)
      logical, intent(in), optional :: l
      integer, intent(in) :: ip
      if (.not. present(l)) return
      if (.not. l) return
      n_keep = n_keep + 1
      keep(n_keep) = ip
    end subroutine flag

  end subroutine set_spectrum

  ! Reads McICA data and attaches it to the named spectra
  ! (socrates_set_spectrum.F90:551-694)
  subroutine set_mcica(mcica_data_file, sw_spectrum_name, lw_spectrum_name)
    character(len=*), intent(in) :: mcica_data_file
    character(len=*), intent(in), optional :: sw_spectrum_name, lw_spectrum_name
    character(kind=c_char, len=256), target :: sw, lw
    type(c_ptr) :: psw, plw
    psw = c_null_ptr
    plw = c_null_ptr
    if (present(sw_spectrum_name)) then
      sw = trim(sw_spectrum_name) // c_null_char
      psw = c_loc(sw)
    end if
    if (present(lw_spectrum_name)) then
      lw = trim(lw_spectrum_name) // c_null_char
      plw = c_loc(lw)
    end if
    call ensure_init()
    call check(socgpu_mcica_register(c_str(mcica_data_file), psw, plw), 'SET_MCICA')
  end subroutine set_mcica

end module socrates_set_spectrum


! Caller-associated outputs (socrates_def_diag.F90:31-175, the path's fields)
module socrates_def_diag
  use realtype_rd, only: RealExt
  implicit none
  private
  public :: StrDiag
  type :: StrDiag
    real(RealExt), pointer :: heating_rate(:, :) => null()
    real(RealExt), pointer :: flux_direct(:, :) => null()
    real(RealExt), pointer :: flux_down(:, :) => null()
    real(RealExt), pointer :: flux_up(:, :) => null()
    real(RealExt), pointer :: flux_direct_clear(:, :) => null()
    real(RealExt), pointer :: flux_down_clear(:, :) => null()
    real(RealExt), pointer :: flux_up_clear(:, :) => null()
    real(RealExt), pointer :: flux_direct_surf(:) => null()
    real(RealExt), pointer :: flux_down_surf(:) => null()
    real(RealExt), pointer :: flux_up_surf(:) => null()
    real(RealExt), pointer :: flux_direct_clear_surf(:) => null()
    real(RealExt), pointer :: flux_down_clear_surf(:) => null()
    real(RealExt), pointer :: flux_up_clear_surf(:) => null()
    real(RealExt), pointer :: flux_direct_toa(:) => null()
    real(RealExt), pointer :: flux_down_toa(:) => null()
    real(RealExt), pointer :: flux_up_toa(:) => null()
    real(RealExt), pointer :: flux_direct_clear_toa(:) => null()
    real(RealExt), pointer :: flux_down_clear_toa(:) => null()
    real(RealExt), pointer :: flux_up_clear_toa(:) => null()
  end type StrDiag
end module socrates_def_diag


module socrates_runes
  use, intrinsic :: iso_c_binding
  use realtype_rd, only: RealK, RealExt
  use socrates_gpu_c
  use def_spectrum, only: StrSpecData
  use socrates_def_diag, only: StrDiag
  implicit none
  private
  public :: runes, StrDiag
  public :: ip_source_illuminate, ip_source_thermal, ip_solar, ip_infra_red
  public :: ip_cloud_off, ip_cloud_homogen, ip_cloud_ice_water, ip_cloud_conv_strat
  public :: ip_max_rand, ip_rand, ip_exponential_rand
  public :: ip_homogeneous, ip_mcica, ip_mcica_single_sampling, ip_mcica_optimal_sampling
  public :: ip_precision_fp64, ip_precision_mixed
  public :: ip_scatter_full, ip_scatter_none, ip_scatter_approx, ip_scatter_hybrid

  ! Option codes (rad_pcf.F90 values)
  integer, parameter :: ip_source_illuminate = 1, ip_source_thermal = 2
  integer, parameter :: ip_solar = 1, ip_infra_red = 2
  integer, parameter :: ip_cloud_off = 0, ip_cloud_homogen = 1
  integer, parameter :: ip_cloud_ice_water = 2, ip_cloud_conv_strat = 3
  integer, parameter :: ip_max_rand = 0, ip_rand = 1, ip_exponential_rand = 3
  integer, parameter :: ip_homogeneous = 0, ip_mcica = 2
  integer, parameter :: ip_mcica_single_sampling = 1, ip_mcica_optimal_sampling = 2
  ! treatment of scattering (socrates_runes.F90:32-35)
  integer, parameter :: ip_scatter_full = 1, ip_scatter_none = 2
  integer, parameter :: ip_scatter_approx = 4, ip_scatter_hybrid = 5
  ! Precision forms of the device solve (no reference counterpart)
  integer, parameter :: ip_precision_fp64 = 0, ip_precision_mixed = 1

  ! library slots of the further gases (SOCGPU_GAS_*) by absorber identifier
  integer, parameter :: n_further = 10
  integer, parameter :: further_ip(n_further) = [4, 6, 7, 9, 25, 14, 15, 16, 17, 19]

contains

  ! One radiation call on the device (socrates_runes.F90:47-1079 on the
  ! two-stream flux path), in either calling form. Outputs through diag as in
  ! the reference. `precision` selects the device form (default FP64, the
  ! reference's).
  subroutine runes(n_profile, n_layer, diag, spectrum, spectrum_name, &
      profile_list, n_layer_stride, n_level_stride, n_cloud_layer, n_subcol_gen, &
      p_layer, t_layer, p_level, t_level, mass, density, &
      h2o, co2, o3, n2o, co, ch4, o2, no, so2, no2, nh3, hno3, n2, cfc11, &
      cfc12, cfc113, hcfc22, hfc125, hfc134a, cfc114, tio, vo, h2, he, ocs, &
      na, k, feh, crh, li, rb, cs, ph3, c2h2, hcn, h2s, ar, o, n, no3, &
      n2o5, hono, ho2no2, h2o2, c2h6, ch3, h2co, ho2, hdo, hcl, hf, cosso, &
      tosso, yosos, ch3cho, ch3ooh, ch3coch3, ch3cocho, chocho, c2h5cho, &
      hoch2cho, c2h5coch3, mvk, macr, pan, ch3ono2, &
      p_layer_1d, t_layer_1d, p_level_1d, t_level_1d, mass_1d, density_1d, &
      h2o_1d, co2_1d, o3_1d, n2o_1d, co_1d, ch4_1d, o2_1d, no_1d, so2_1d, &
      no2_1d, nh3_1d, hno3_1d, n2_1d, cfc11_1d, cfc12_1d, cfc113_1d, &
      hcfc22_1d, hfc125_1d, hfc134a_1d, cfc114_1d, tio_1d, vo_1d, h2_1d, &
      he_1d, ocs_1d, na_1d, k_1d, feh_1d, crh_1d, li_1d, rb_1d, cs_1d, &
      ph3_1d, c2h2_1d, hcn_1d, h2s_1d, ar_1d, o_1d, n_1d, no3_1d, n2o5_1d, &
      hono_1d, ho2no2_1d, h2o2_1d, c2h6_1d, ch3_1d, h2co_1d, ho2_1d, &
      hdo_1d, hcl_1d, hf_1d, cosso_1d, tosso_1d, yosos_1d, ch3cho_1d, &
      ch3ooh_1d, ch3coch3_1d, ch3cocho_1d, chocho_1d, c2h5cho_1d, &
      hoch2cho_1d, c2h5coch3_1d, mvk_1d, macr_1d, pan_1d, ch3ono2_1d, &
      h2o_mix_ratio, co2_mix_ratio, o3_mix_ratio, n2o_mix_ratio, &
      co_mix_ratio, ch4_mix_ratio, o2_mix_ratio, no_mix_ratio, &
      so2_mix_ratio, no2_mix_ratio, nh3_mix_ratio, hno3_mix_ratio, &
      n2_mix_ratio, cfc11_mix_ratio, cfc12_mix_ratio, cfc113_mix_ratio, &
      hcfc22_mix_ratio, hfc125_mix_ratio, hfc134a_mix_ratio, &
      cfc114_mix_ratio, tio_mix_ratio, vo_mix_ratio, h2_mix_ratio, &
      he_mix_ratio, ocs_mix_ratio, na_mix_ratio, k_mix_ratio, &
      feh_mix_ratio, crh_mix_ratio, li_mix_ratio, rb_mix_ratio, &
      cs_mix_ratio, ph3_mix_ratio, c2h2_mix_ratio, hcn_mix_ratio, &
      h2s_mix_ratio, ar_mix_ratio, o_mix_ratio, n_mix_ratio, no3_mix_ratio, &
      n2o5_mix_ratio, hono_mix_ratio, ho2no2_mix_ratio, h2o2_mix_ratio, &
      c2h6_mix_ratio, ch3_mix_ratio, h2co_mix_ratio, ho2_mix_ratio, &
      hdo_mix_ratio, hcl_mix_ratio, hf_mix_ratio, cosso_mix_ratio, &
      tosso_mix_ratio, yosos_mix_ratio, ch3cho_mix_ratio, ch3ooh_mix_ratio, &
      ch3coch3_mix_ratio, ch3cocho_mix_ratio, chocho_mix_ratio, &
      c2h5cho_mix_ratio, hoch2cho_mix_ratio, c2h5coch3_mix_ratio, &
      mvk_mix_ratio, macr_mix_ratio, pan_mix_ratio, ch3ono2_mix_ratio, &
      l_h2o_well_mixed, l_co2_well_mixed, l_o3_well_mixed, &
      l_n2o_well_mixed, l_co_well_mixed, l_ch4_well_mixed, l_o2_well_mixed, &
      l_no_well_mixed, l_so2_well_mixed, l_no2_well_mixed, &
      l_nh3_well_mixed, l_hno3_well_mixed, l_n2_well_mixed, &
      l_cfc11_well_mixed, l_cfc12_well_mixed, l_cfc113_well_mixed, &
      l_hcfc22_well_mixed, l_hfc125_well_mixed, l_hfc134a_well_mixed, &
      l_cfc114_well_mixed, l_tio_well_mixed, l_vo_well_mixed, &
      l_h2_well_mixed, l_he_well_mixed, l_ocs_well_mixed, l_na_well_mixed, &
      l_k_well_mixed, l_feh_well_mixed, l_crh_well_mixed, l_li_well_mixed, &
      l_rb_well_mixed, l_cs_well_mixed, l_ph3_well_mixed, &
      l_c2h2_well_mixed, l_hcn_well_mixed, l_h2s_well_mixed, &
      l_ar_well_mixed, l_o_well_mixed, l_n_well_mixed, l_no3_well_mixed, &
      l_n2o5_well_mixed, l_hono_well_mixed, l_ho2no2_well_mixed, &
      l_h2o2_well_mixed, l_c2h6_well_mixed, l_ch3_well_mixed, &
      l_h2co_well_mixed, l_ho2_well_mixed, l_hdo_well_mixed, &
      l_hcl_well_mixed, l_hf_well_mixed, l_cosso_well_mixed, &
      l_tosso_well_mixed, l_yosos_well_mixed, l_ch3cho_well_mixed, &
      l_ch3ooh_well_mixed, l_ch3coch3_well_mixed, l_ch3cocho_well_mixed, &
      l_chocho_well_mixed, l_c2h5cho_well_mixed, l_hoch2cho_well_mixed, &
      l_c2h5coch3_well_mixed, l_mvk_well_mixed, l_macr_well_mixed, &
      l_pan_well_mixed, l_ch3ono2_well_mixed, &
      l_flux_ground, t_ground, flux_ground, flux_ground_1d, &
      cos_zenith_angle, solar_irrad, &
      l_grey_albedo, grey_albedo, albedo_diff, albedo_dir, &
      albedo_diff_1d, albedo_dir_1d, &
      cloud_frac, conv_frac, liq_mmr, ice_mmr, liq_conv_mmr, ice_conv_mmr, &
      liq_dim, ice_dim, liq_conv_dim, ice_conv_dim, &
      cloud_frac_1d, conv_frac_1d, liq_mmr_1d, ice_mmr_1d, liq_conv_mmr_1d, &
      ice_conv_mmr_1d, liq_dim_1d, ice_dim_1d, liq_conv_dim_1d, ice_conv_dim_1d, &
      cloud_vertical_decorr, conv_vertical_decorr, cloud_horizontal_rsd, &
      layer_heat_capacity, layer_heat_capacity_1d, &
      i_source, i_scatter_method, i_cloud_representation, i_overlap, i_inhom, &
      i_mcica_sampling, &
      i_st_water, i_cnv_water, i_st_ice, i_cnv_ice, rand_seed, &
      n_aer_mode, n_aer_layer, l_aerosol_mode, aer_mix_ratio, aer_absorption, &
      aer_scattering, aer_asymmetry, aer_mix_ratio_1d, aer_absorption_1d, &
      aer_scattering_1d, aer_asymmetry_1d, &
      l_rayleigh, l_invert, l_profile_last, precision)
    integer, intent(in) :: n_profile, n_layer
    type(StrDiag), intent(inout) :: diag
    type(StrSpecData), intent(in), optional :: spectrum
    character(len=*), intent(in), optional :: spectrum_name
    integer, intent(in), optional :: profile_list(:)
    integer, intent(in), optional :: n_layer_stride, n_level_stride, n_cloud_layer
    integer, intent(in), optional :: n_subcol_gen
    real(RealExt), intent(in), optional :: p_layer(:, :), t_layer(:, :)
    real(RealExt), intent(in), optional :: p_level(:, :), t_level(:, :)
    real(RealExt), intent(in), optional :: mass(:, :), density(:, :)
    real(RealExt), intent(in), optional :: &
      h2o(:, :), co2(:, :), o3(:, :), n2o(:, :), co(:, :), ch4(:, :), &
      o2(:, :), no(:, :), so2(:, :), no2(:, :), nh3(:, :), hno3(:, :), &
      n2(:, :), cfc11(:, :), cfc12(:, :), cfc113(:, :), hcfc22(:, :), &
      hfc125(:, :), hfc134a(:, :), cfc114(:, :), tio(:, :), vo(:, :), &
      h2(:, :), he(:, :), ocs(:, :), na(:, :), k(:, :), feh(:, :), &
      crh(:, :), li(:, :), rb(:, :), cs(:, :), ph3(:, :), c2h2(:, :), &
      hcn(:, :), h2s(:, :), ar(:, :), o(:, :), n(:, :), no3(:, :), &
      n2o5(:, :), hono(:, :), ho2no2(:, :), h2o2(:, :), c2h6(:, :), &
      ch3(:, :), h2co(:, :), ho2(:, :), hdo(:, :), hcl(:, :), hf(:, :), &
      cosso(:, :), tosso(:, :), yosos(:, :), ch3cho(:, :), ch3ooh(:, :), &
      ch3coch3(:, :), ch3cocho(:, :), chocho(:, :), c2h5cho(:, :), &
      hoch2cho(:, :), c2h5coch3(:, :), mvk(:, :), macr(:, :), pan(:, :), &
      ch3ono2(:, :)
    real(RealExt), intent(in), optional :: p_layer_1d(:), t_layer_1d(:)
    real(RealExt), intent(in), optional :: p_level_1d(:), t_level_1d(:)
    real(RealExt), intent(in), optional :: mass_1d(:), density_1d(:)
    real(RealExt), intent(in), optional :: &
      h2o_1d(:), co2_1d(:), o3_1d(:), n2o_1d(:), co_1d(:), ch4_1d(:), &
      o2_1d(:), no_1d(:), so2_1d(:), no2_1d(:), nh3_1d(:), hno3_1d(:), &
      n2_1d(:), cfc11_1d(:), cfc12_1d(:), cfc113_1d(:), hcfc22_1d(:), &
      hfc125_1d(:), hfc134a_1d(:), cfc114_1d(:), tio_1d(:), vo_1d(:), &
      h2_1d(:), he_1d(:), ocs_1d(:), na_1d(:), k_1d(:), feh_1d(:), &
      crh_1d(:), li_1d(:), rb_1d(:), cs_1d(:), ph3_1d(:), c2h2_1d(:), &
      hcn_1d(:), h2s_1d(:), ar_1d(:), o_1d(:), n_1d(:), no3_1d(:), &
      n2o5_1d(:), hono_1d(:), ho2no2_1d(:), h2o2_1d(:), c2h6_1d(:), &
      ch3_1d(:), h2co_1d(:), ho2_1d(:), hdo_1d(:), hcl_1d(:), hf_1d(:), &
      cosso_1d(:), tosso_1d(:), yosos_1d(:), ch3cho_1d(:), ch3ooh_1d(:), &
      ch3coch3_1d(:), ch3cocho_1d(:), chocho_1d(:), c2h5cho_1d(:), &
      hoch2cho_1d(:), c2h5coch3_1d(:), mvk_1d(:), macr_1d(:), pan_1d(:), &
      ch3ono2_1d(:)
    real(RealExt), intent(in), optional :: &
      h2o_mix_ratio, co2_mix_ratio, o3_mix_ratio, n2o_mix_ratio, &
      co_mix_ratio, ch4_mix_ratio, o2_mix_ratio, no_mix_ratio, &
      so2_mix_ratio, no2_mix_ratio, nh3_mix_ratio, hno3_mix_ratio, &
      n2_mix_ratio, cfc11_mix_ratio, cfc12_mix_ratio, cfc113_mix_ratio, &
      hcfc22_mix_ratio, hfc125_mix_ratio, hfc134a_mix_ratio, &
      cfc114_mix_ratio, tio_mix_ratio, vo_mix_ratio, h2_mix_ratio, &
      he_mix_ratio, ocs_mix_ratio, na_mix_ratio, k_mix_ratio, &
      feh_mix_ratio, crh_mix_ratio, li_mix_ratio, rb_mix_ratio, &
      cs_mix_ratio, ph3_mix_ratio, c2h2_mix_ratio, hcn_mix_ratio, &
      h2s_mix_ratio, ar_mix_ratio, o_mix_ratio, n_mix_ratio, no3_mix_ratio, &
      n2o5_mix_ratio, hono_mix_ratio, ho2no2_mix_ratio, h2o2_mix_ratio, &
      c2h6_mix_ratio, ch3_mix_ratio, h2co_mix_ratio, ho2_mix_ratio, &
      hdo_mix_ratio, hcl_mix_ratio, hf_mix_ratio, cosso_mix_ratio, &
      tosso_mix_ratio, yosos_mix_ratio, ch3cho_mix_ratio, ch3ooh_mix_ratio, &
      ch3coch3_mix_ratio, ch3cocho_mix_ratio, chocho_mix_ratio, &
      c2h5cho_mix_ratio, hoch2cho_mix_ratio, c2h5coch3_mix_ratio, &
      mvk_mix_ratio, macr_mix_ratio, pan_mix_ratio, ch3ono2_mix_ratio
    logical, intent(in), optional :: &
      l_h2o_well_mixed, l_co2_well_mixed, l_o3_well_mixed, &
      l_n2o_well_mixed, l_co_well_mixed, l_ch4_well_mixed, l_o2_well_mixed, &
      l_no_well_mixed, l_so2_well_mixed, l_no2_well_mixed, &
      l_nh3_well_mixed, l_hno3_well_mixed, l_n2_well_mixed, &
      l_cfc11_well_mixed, l_cfc12_well_mixed, l_cfc113_well_mixed, &
      l_hcfc22_well_mixed, l_hfc125_well_mixed, l_hfc134a_well_mixed, &
      l_cfc114_well_mixed, l_tio_well_mixed, l_vo_well_mixed, &
      l_h2_well_mixed, l_he_well_mixed, l_ocs_well_mixed, l_na_well_mixed, &
      l_k_well_mixed, l_feh_well_mixed, l_crh_well_mixed, l_li_well_mixed, &
      l_rb_well_mixed, l_cs_well_mixed, l_ph3_well_mixed, &
      l_c2h2_well_mixed, l_hcn_well_mixed, l_h2s_well_mixed, &
      l_ar_well_mixed, l_o_well_mixed, l_n_well_mixed, l_no3_well_mixed, &
      l_n2o5_well_mixed, l_hono_well_mixed, l_ho2no2_well_mixed, &
      l_h2o2_well_mixed, l_c2h6_well_mixed, l_ch3_well_mixed, &
      l_h2co_well_mixed, l_ho2_well_mixed, l_hdo_well_mixed, &
      l_hcl_well_mixed, l_hf_well_mixed, l_cosso_well_mixed, &
      l_tosso_well_mixed, l_yosos_well_mixed, l_ch3cho_well_mixed, &
      l_ch3ooh_well_mixed, l_ch3coch3_well_mixed, l_ch3cocho_well_mixed, &
      l_chocho_well_mixed, l_c2h5cho_well_mixed, l_hoch2cho_well_mixed, &
      l_c2h5coch3_well_mixed, l_mvk_well_mixed, l_macr_well_mixed, &
      l_pan_well_mixed, l_ch3ono2_well_mixed
    logical, intent(in), optional :: l_flux_ground
    real(RealExt), intent(in), optional :: t_ground(:)
    real(RealExt), intent(in), optional :: flux_ground(:, :), flux_ground_1d(:)
    real(RealExt), intent(in), optional :: cos_zenith_angle(:), solar_irrad(:)
    logical, intent(in), optional :: l_grey_albedo
    real(RealExt), intent(in), optional :: grey_albedo
    real(RealExt), intent(in), optional :: albedo_diff(:, :), albedo_dir(:, :)
    real(RealExt), intent(in), optional :: albedo_diff_1d(:), albedo_dir_1d(:)
    real(RealExt), intent(in), optional :: cloud_frac(:, :), conv_frac(:, :)
    real(RealExt), intent(in), optional :: liq_mmr(:, :), ice_mmr(:, :)
    real(RealExt), intent(in), optional :: liq_conv_mmr(:, :), ice_conv_mmr(:, :)
    real(RealExt), intent(in), optional :: liq_dim(:, :), ice_dim(:, :)
    real(RealExt), intent(in), optional :: liq_conv_dim(:, :), ice_conv_dim(:, :)
    real(RealExt), intent(in), optional :: cloud_frac_1d(:), conv_frac_1d(:)
    real(RealExt), intent(in), optional :: liq_mmr_1d(:), ice_mmr_1d(:)
    real(RealExt), intent(in), optional :: liq_conv_mmr_1d(:), ice_conv_mmr_1d(:)
    real(RealExt), intent(in), optional :: liq_dim_1d(:), ice_dim_1d(:)
    real(RealExt), intent(in), optional :: liq_conv_dim_1d(:), ice_conv_dim_1d(:)
    real(RealExt), intent(in), optional :: cloud_vertical_decorr, conv_vertical_decorr
    real(RealExt), intent(in), optional :: cloud_horizontal_rsd
    real(RealExt), intent(in), optional :: layer_heat_capacity(:, :)
    real(RealExt), intent(in), optional :: layer_heat_capacity_1d(:)
    integer, intent(in), optional :: i_source, i_cloud_representation, i_overlap, i_inhom
    integer, intent(in), optional :: i_scatter_method
    integer, intent(in), optional :: i_mcica_sampling, i_st_water, i_cnv_water
    integer, intent(in), optional :: i_st_ice, i_cnv_ice
    integer, intent(in), optional :: rand_seed(:)
    integer, intent(in), optional :: n_aer_mode, n_aer_layer
    logical, intent(in), optional :: l_aerosol_mode
    real(RealExt), intent(in), optional :: aer_mix_ratio(:, :, :)
    real(RealExt), intent(in), optional :: aer_absorption(:, :, :, :), &
      aer_scattering(:, :, :, :), aer_asymmetry(:, :, :, :)
    real(RealExt), intent(in), optional :: aer_mix_ratio_1d(:), aer_absorption_1d(:), &
      aer_scattering_1d(:), aer_asymmetry_1d(:)
    logical, intent(in), optional :: l_rayleigh, l_invert, l_profile_last
    integer, intent(in), optional :: precision

    ! dense copies in the library's layout: (n_profile, layers or levels or
    ! bands), layer 1 at the top
    type :: buf
      real(RealExt), allocatable :: v(:, :)
    end type buf
    integer, parameter :: n_buf = 56
    type(buf), target :: b(n_buf)
    integer :: nb_used
    type(socgpu_runes_args) :: a
    type(socgpu_diag) :: d
    integer :: list(n_profile), l, i, ii, ll, id_ct
    integer :: layer_offset, level_offset, stride_layer, stride_level, stride_aer, n_mode
    logical :: last, cloud
    integer(c_int) :: prec, nb, nabs, types(80), status
    integer(c_int), allocatable, target :: seeds(:)
    character(len=64) :: name
    integer :: o_up, o_dn, o_dir, o_hr, o_upc, o_dnc, o_dirc

    call ensure_init()
    nb_used = 0
    if (present(spectrum)) then
      if (spectrum%name == '') call fail('spectrum has not been set (set_spectrum)')
      name = spectrum%name
    else if (present(spectrum_name)) then
      name = spectrum_name
    else
      call fail('spectrum or spectrum_name is required')
    end if
    call check(socgpu_spectrum_query(c_str(name), nb, nabs, types, 80_c_int), 'RUNES')

    ! profile list, vertical order and strides (socrates_set_atm.F90:191-225)
    if (present(profile_list)) then
      if (size(profile_list) < n_profile) call fail('profile_list has fewer than n_profile entries')
      list = profile_list(1:n_profile)
    else
      list = [(l, l = 1, n_profile)]
    end if
    layer_offset = 0
    level_offset = 0
    if (present(l_invert)) then
      if (l_invert) then
        layer_offset = n_layer + 1
        level_offset = n_layer
      end if
    end if
    last = .false.
    if (present(l_profile_last)) last = l_profile_last
    stride_layer = n_layer
    if (present(n_layer_stride)) stride_layer = n_layer_stride
    stride_level = n_layer + 1
    if (present(n_level_stride)) stride_level = n_level_stride

    a%n_profile = n_profile
    a%n_layer = n_layer
    if (present(i_source)) a%i_source = i_source
    if (present(l_rayleigh)) a%l_rayleigh = merge(1, 0, l_rayleigh)
    if (present(i_cloud_representation)) a%i_cloud_representation = i_cloud_representation
    if (present(i_overlap)) a%i_overlap = i_overlap
    if (present(i_scatter_method)) a%i_scatter_method = i_scatter_method
    if (present(i_inhom)) a%i_inhom = i_inhom
    if (present(i_mcica_sampling)) a%i_mcica_sampling = i_mcica_sampling
    if (present(i_st_water)) a%i_st_water = i_st_water
    if (present(i_cnv_water)) a%i_cnv_water = i_cnv_water
    if (present(n_subcol_gen)) a%n_subcol_gen = n_subcol_gen
    if (present(cloud_vertical_decorr)) a%cloud_vertical_decorr = cloud_vertical_decorr
    if (present(conv_vertical_decorr)) a%conv_vertical_decorr = conv_vertical_decorr
    if (present(cloud_horizontal_rsd)) a%cloud_horizontal_rsd = cloud_horizontal_rsd
    ! grey albedo only with its value (socrates_set_bound.F90:138)
    if (present(l_grey_albedo) .and. present(grey_albedo)) then
      a%l_grey_albedo = merge(1, 0, l_grey_albedo)
      a%grey_albedo = grey_albedo
    end if

    a%p_layer = lay(p_layer, p_layer_1d)
    a%t_layer = lay(t_layer, t_layer_1d)
    a%mass = lay(mass, mass_1d)
    a%density = lay(density, density_1d)
    a%layer_heat_capacity = lay(layer_heat_capacity, layer_heat_capacity_1d)
    a%p_level = lev(p_level, p_level_1d)
    a%t_level = lev(t_level, t_level_1d)
    a%albedo_diff = band(albedo_diff, albedo_diff_1d)
    a%albedo_dir = band(albedo_dir, albedo_dir_1d)
    a%t_ground = col(t_ground)
    a%cos_zenith_angle = col(cos_zenith_angle)
    a%solar_irrad = col(solar_irrad)
    if (present(l_flux_ground)) then
      if (l_flux_ground) then
        a%l_flux_ground = 1
        a%flux_ground = band(flux_ground, flux_ground_1d)
      end if
    end if

    ! gases: a field, a strided field or a well-mixed ratio each, for the
    ! absorbers the spectral file lists (socrates_set_atm.F90:246-300, 535-622)
    call gas(1, 'h2o', h2o, h2o_1d, h2o_mix_ratio, l_h2o_well_mixed)
    call gas(2, 'co2', co2, co2_1d, co2_mix_ratio, l_co2_well_mixed)
    call gas(3, 'o3', o3, o3_1d, o3_mix_ratio, l_o3_well_mixed)
    call gas(4, 'n2o', n2o, n2o_1d, n2o_mix_ratio, l_n2o_well_mixed)
    call gas(5, 'co', co, co_1d, co_mix_ratio, l_co_well_mixed)
    call gas(6, 'ch4', ch4, ch4_1d, ch4_mix_ratio, l_ch4_well_mixed)
    call gas(7, 'o2', o2, o2_1d, o2_mix_ratio, l_o2_well_mixed)
    call gas(8, 'no', no, no_1d, no_mix_ratio, l_no_well_mixed)
    call gas(9, 'so2', so2, so2_1d, so2_mix_ratio, l_so2_well_mixed)
    call gas(10, 'no2', no2, no2_1d, no2_mix_ratio, l_no2_well_mixed)
    call gas(11, 'nh3', nh3, nh3_1d, nh3_mix_ratio, l_nh3_well_mixed)
    call gas(12, 'hno3', hno3, hno3_1d, hno3_mix_ratio, l_hno3_well_mixed)
    call gas(13, 'n2', n2, n2_1d, n2_mix_ratio, l_n2_well_mixed)
    call gas(14, 'cfc11', cfc11, cfc11_1d, cfc11_mix_ratio, l_cfc11_well_mixed)
    call gas(15, 'cfc12', cfc12, cfc12_1d, cfc12_mix_ratio, l_cfc12_well_mixed)
    call gas(16, 'cfc113', cfc113, cfc113_1d, cfc113_mix_ratio, l_cfc113_well_mixed)
    call gas(17, 'hcfc22', hcfc22, hcfc22_1d, hcfc22_mix_ratio, l_hcfc22_well_mixed)
    call gas(18, 'hfc125', hfc125, hfc125_1d, hfc125_mix_ratio, l_hfc125_well_mixed)
    call gas(19, 'hfc134a', hfc134a, hfc134a_1d, hfc134a_mix_ratio, l_hfc134a_well_mixed)
    call gas(20, 'cfc114', cfc114, cfc114_1d, cfc114_mix_ratio, l_cfc114_well_mixed)
    call gas(21, 'tio', tio, tio_1d, tio_mix_ratio, l_tio_well_mixed)
    call gas(22, 'vo', vo, vo_1d, vo_mix_ratio, l_vo_well_mixed)
    call gas(23, 'h2', h2, h2_1d, h2_mix_ratio, l_h2_well_mixed)
    call gas(24, 'he', he, he_1d, he_mix_ratio, l_he_well_mixed)
    call gas(25, 'ocs', ocs, ocs_1d, ocs_mix_ratio, l_ocs_well_mixed)
    call gas(26, 'na', na, na_1d, na_mix_ratio, l_na_well_mixed)
    call gas(27, 'k', k, k_1d, k_mix_ratio, l_k_well_mixed)
    call gas(28, 'feh', feh, feh_1d, feh_mix_ratio, l_feh_well_mixed)
    call gas(29, 'crh', crh, crh_1d, crh_mix_ratio, l_crh_well_mixed)
    call gas(30, 'li', li, li_1d, li_mix_ratio, l_li_well_mixed)
    call gas(31, 'rb', rb, rb_1d, rb_mix_ratio, l_rb_well_mixed)
    call gas(32, 'cs', cs, cs_1d, cs_mix_ratio, l_cs_well_mixed)
    call gas(33, 'ph3', ph3, ph3_1d, ph3_mix_ratio, l_ph3_well_mixed)
    call gas(34, 'c2h2', c2h2, c2h2_1d, c2h2_mix_ratio, l_c2h2_well_mixed)
    call gas(35, 'hcn', hcn, hcn_1d, hcn_mix_ratio, l_hcn_well_mixed)
    call gas(36, 'h2s', h2s, h2s_1d, h2s_mix_ratio, l_h2s_well_mixed)
    call gas(37, 'ar', ar, ar_1d, ar_mix_ratio, l_ar_well_mixed)
    call gas(39, 'o', o, o_1d, o_mix_ratio, l_o_well_mixed)
    call gas(40, 'n', n, n_1d, n_mix_ratio, l_n_well_mixed)
    call gas(41, 'no3', no3, no3_1d, no3_mix_ratio, l_no3_well_mixed)
    call gas(42, 'n2o5', n2o5, n2o5_1d, n2o5_mix_ratio, l_n2o5_well_mixed)
    call gas(43, 'hono', hono, hono_1d, hono_mix_ratio, l_hono_well_mixed)
    call gas(44, 'ho2no2', ho2no2, ho2no2_1d, ho2no2_mix_ratio, l_ho2no2_well_mixed)
    call gas(45, 'h2o2', h2o2, h2o2_1d, h2o2_mix_ratio, l_h2o2_well_mixed)
    call gas(46, 'c2h6', c2h6, c2h6_1d, c2h6_mix_ratio, l_c2h6_well_mixed)
    call gas(47, 'ch3', ch3, ch3_1d, ch3_mix_ratio, l_ch3_well_mixed)
    call gas(48, 'h2co', h2co, h2co_1d, h2co_mix_ratio, l_h2co_well_mixed)
    call gas(49, 'ho2', ho2, ho2_1d, ho2_mix_ratio, l_ho2_well_mixed)
    call gas(50, 'hdo', hdo, hdo_1d, hdo_mix_ratio, l_hdo_well_mixed)
    call gas(51, 'hcl', hcl, hcl_1d, hcl_mix_ratio, l_hcl_well_mixed)
    call gas(52, 'hf', hf, hf_1d, hf_mix_ratio, l_hf_well_mixed)
    call gas(53, 'cosso', cosso, cosso_1d, cosso_mix_ratio, l_cosso_well_mixed)
    call gas(54, 'tosso', tosso, tosso_1d, tosso_mix_ratio, l_tosso_well_mixed)
    call gas(55, 'yosos', yosos, yosos_1d, yosos_mix_ratio, l_yosos_well_mixed)
    call gas(56, 'ch3cho', ch3cho, ch3cho_1d, ch3cho_mix_ratio, l_ch3cho_well_mixed)
    call gas(57, 'ch3ooh', ch3ooh, ch3ooh_1d, ch3ooh_mix_ratio, l_ch3ooh_well_mixed)
    call gas(58, 'ch3coch3', ch3coch3, ch3coch3_1d, ch3coch3_mix_ratio, l_ch3coch3_well_mixed)
    call gas(59, 'ch3cocho', ch3cocho, ch3cocho_1d, ch3cocho_mix_ratio, l_ch3cocho_well_mixed)
    call gas(60, 'chocho', chocho, chocho_1d, chocho_mix_ratio, l_chocho_well_mixed)
    call gas(61, 'c2h5cho', c2h5cho, c2h5cho_1d, c2h5cho_mix_ratio, l_c2h5cho_well_mixed)
    call gas(62, 'hoch2cho', hoch2cho, hoch2cho_1d, hoch2cho_mix_ratio, l_hoch2cho_well_mixed)
    call gas(63, 'c2h5coch3', c2h5coch3, c2h5coch3_1d, c2h5coch3_mix_ratio, l_c2h5coch3_well_mixed)
    call gas(64, 'mvk', mvk, mvk_1d, mvk_mix_ratio, l_mvk_well_mixed)
    call gas(65, 'macr', macr, macr_1d, macr_mix_ratio, l_macr_well_mixed)
    call gas(66, 'pan', pan, pan_1d, pan_mix_ratio, l_pan_well_mixed)
    call gas(67, 'ch3ono2', ch3ono2, ch3ono2_1d, ch3ono2_mix_ratio, l_ch3ono2_well_mixed)

    ! cloud (socrates_set_control.F90:296-301: no cloudy layers is cloud off;
    ! socrates_set_dimen.F90:121-123 and socrates_set_cld.F90:861-910: cloud
    ! only from layer n_layer + 1 - n_cloud_layer down)
    cloud = a%i_cloud_representation /= ip_cloud_off
    id_ct = 1
    if (present(n_cloud_layer)) then
      if (n_cloud_layer < 1) then
        a%i_cloud_representation = ip_cloud_off
        cloud = .false.
      end if
      id_ct = n_layer + 1 - max(1, n_cloud_layer)
    end if
    if (cloud) then
      a%cloud_frac = lay(cloud_frac, cloud_frac_1d, id_ct)
      a%liq_mmr = lay(liq_mmr, liq_mmr_1d, id_ct)
      a%liq_dim = lay(liq_dim, liq_dim_1d, id_ct)
      a%conv_frac = lay(conv_frac, conv_frac_1d, id_ct)
      a%liq_conv_mmr = lay(liq_conv_mmr, liq_conv_mmr_1d, id_ct)
      a%liq_conv_dim = lay(liq_conv_dim, liq_conv_dim_1d, id_ct)
      ! ice (socrates_set_cld.F90:341-415)
      if (present(i_st_ice)) a%i_st_ice = i_st_ice
      if (present(i_cnv_ice)) a%i_cnv_ice = i_cnv_ice
      a%ice_mmr = lay(ice_mmr, ice_mmr_1d, id_ct)
      a%ice_dim = lay(ice_dim, ice_dim_1d, id_ct)
      a%ice_conv_mmr = lay(ice_conv_mmr, ice_conv_mmr_1d, id_ct)
      a%ice_conv_dim = lay(ice_conv_dim, ice_conv_dim_1d, id_ct)
      ! McICA seeds (socrates_set_cld_mcica.F90:159-167)
      allocate(seeds(n_profile))
      if (present(rand_seed)) then
        seeds = int(rand_seed(list), c_int)
      else
        seeds = int([(10 + l, l = 1, n_profile)], c_int)
      end if
      a%rand_seed = c_loc(seeds)
    end if

    ! modal aerosols (socrates_set_control.F90:378-384: on when modes are
    ! given, unless l_aerosol_mode is false; socrates_set_aer.F90:299-309,
    ! 433-566: an absent field is zero)
    n_mode = 0
    if (present(n_aer_mode)) n_mode = max(n_aer_mode, 0)
    if (present(l_aerosol_mode)) then
      if (.not. l_aerosol_mode) n_mode = 0
    end if
    if (n_mode > 0) then
      stride_aer = stride_layer
      if (present(n_aer_layer)) stride_aer = n_aer_layer
      a%n_aer_mode = n_mode
      a%aer_mix_ratio = mode_field(aer_mix_ratio, aer_mix_ratio_1d)
      a%aer_absorption = mode_prop(aer_absorption, aer_absorption_1d)
      a%aer_scattering = mode_prop(aer_scattering, aer_scattering_1d)
      a%aer_asymmetry = mode_prop(aer_asymmetry, aer_asymmetry_1d)
    end if

    ! outputs: dense buffers, scattered as set_diag places them
    o_up = 0; o_dn = 0; o_dir = 0; o_hr = 0; o_upc = 0; o_dnc = 0; o_dirc = 0
    if (associated(diag%flux_up) .or. associated(diag%flux_up_surf) .or. &
        associated(diag%flux_up_toa)) o_up = out(d%flux_up, n_layer + 1)
    if (associated(diag%flux_down) .or. associated(diag%flux_down_surf) .or. &
        associated(diag%flux_down_toa)) o_dn = out(d%flux_down, n_layer + 1)
    if (associated(diag%flux_direct) .or. associated(diag%flux_direct_surf) .or. &
        associated(diag%flux_direct_toa)) o_dir = out(d%flux_direct, n_layer + 1)
    if (associated(diag%heating_rate)) o_hr = out(d%heating_rate, n_layer)
    if (associated(diag%flux_up_clear) .or. associated(diag%flux_up_clear_surf) .or. &
        associated(diag%flux_up_clear_toa)) o_upc = out(d%flux_up_clear, n_layer + 1)
    if (associated(diag%flux_down_clear) .or. associated(diag%flux_down_clear_surf) .or. &
        associated(diag%flux_down_clear_toa)) o_dnc = out(d%flux_down_clear, n_layer + 1)
    if (associated(diag%flux_direct_clear) .or. associated(diag%flux_direct_clear_surf) .or. &
        associated(diag%flux_direct_clear_toa)) o_dirc = out(d%flux_direct_clear, n_layer + 1)

    prec = ip_precision_fp64
    if (present(precision)) prec = precision
    status = socgpu_runes(c_str(name), a, d, prec)
    call check(status, 'RUNES')

    call give_flux(diag%flux_up, diag%flux_up_surf, diag%flux_up_toa, o_up)
    call give_flux(diag%flux_down, diag%flux_down_surf, diag%flux_down_toa, o_dn)
    call give_flux(diag%flux_direct, diag%flux_direct_surf, diag%flux_direct_toa, o_dir)
    call give_flux(diag%flux_up_clear, diag%flux_up_clear_surf, diag%flux_up_clear_toa, o_upc)
    call give_flux(diag%flux_down_clear, diag%flux_down_clear_surf, &
                   diag%flux_down_clear_toa, o_dnc)
    call give_flux(diag%flux_direct_clear, diag%flux_direct_clear_surf, &
                   diag%flux_direct_clear_toa, o_dirc)
    if (o_hr > 0) call give_heating(diag%heating_rate, o_hr)

  contains

    subroutine fail(msg)
      character(len=*), intent(in) :: msg
      write(*, '(a, a)') '? Error from routine: RUNES: ', msg
      error stop 1
    end subroutine fail

    ! a new dense buffer of n2 rows per profile
    integer function new_buf(n2)
      integer, intent(in) :: n2
      nb_used = nb_used + 1
      if (nb_used > n_buf) call fail('too many fields')
      allocate(b(nb_used)%v(n_profile, n2))
      new_buf = nb_used
    end function new_buf

    ! a layer field from its full or strided form (set_layer_field,
    ! socrates_set_atm.F90:439-482), layer 1 at the top; layers above k0
    ! (cloud top of the call) zero
    type(c_ptr) function lay(full, oned, k0) result(p)
      real(RealExt), intent(in), optional :: full(:, :), oned(:)
      integer, intent(in), optional :: k0
      integer :: j, top
      p = c_null_ptr
      if (.not. (present(full) .or. present(oned))) return
      j = new_buf(n_layer)
      top = 1
      if (present(k0)) top = k0
      b(j)%v = 0.0_RealExt
      do i = top, n_layer
        ii = abs(layer_offset - i)
        do l = 1, n_profile
          if (present(full)) then
            if (last) then
              b(j)%v(l, i) = full(ii, list(l))
            else
              b(j)%v(l, i) = full(list(l), ii)
            end if
          else
            if (last) then
              ll = stride_layer * (list(l) - 1) + ii
            else
              ll = n_profile * (ii - 1) + list(l)
            end if
            b(j)%v(l, i) = oned(ll)
          end if
        end do
      end do
      p = c_loc(b(j)%v)
    end function lay

    ! a level field (set_level_field, socrates_set_atm.F90:485-532), level 0
    ! at the top
    type(c_ptr) function lev(full, oned) result(p)
      real(RealExt), intent(in), optional :: full(:, :), oned(:)
      integer :: j
      p = c_null_ptr
      if (.not. (present(full) .or. present(oned))) return
      j = new_buf(n_layer + 1)
      do i = 0, n_layer
        ii = abs(level_offset - i) + 1
        do l = 1, n_profile
          if (present(full)) then
            if (last) then
              b(j)%v(l, i + 1) = full(ii, list(l))
            else
              b(j)%v(l, i + 1) = full(list(l), ii)
            end if
          else
            if (last) then
              ll = stride_level * (list(l) - 1) + ii
            else
              ll = n_profile * (ii - 1) + list(l)
            end if
            b(j)%v(l, i + 1) = oned(ll)
          end if
        end do
      end do
      p = c_loc(b(j)%v)
    end function lev

    ! a per-band surface field (socrates_set_bound.F90:154-220, 251-290)
    type(c_ptr) function band(full, oned) result(p)
      real(RealExt), intent(in), optional :: full(:, :), oned(:)
      integer :: j, ib
      p = c_null_ptr
      if (.not. (present(full) .or. present(oned))) return
      j = new_buf(int(nb))
      do ib = 1, nb
        do l = 1, n_profile
          if (present(full)) then
            if (last) then
              b(j)%v(l, ib) = full(ib, list(l))
            else
              b(j)%v(l, ib) = full(list(l), ib)
            end if
          else
            if (last) then
              ll = nb * (list(l) - 1) + ib
            else
              ll = n_profile * (ib - 1) + list(l)
            end if
            b(j)%v(l, ib) = oned(ll)
          end if
        end do
      end do
      p = c_loc(b(j)%v)
    end function band

    ! a modal aerosol mixing ratio (set_mode_field,
    ! socrates_set_aer.F90:433-491): buffer (profile, layer + n_layer*(mode-1))
    type(c_ptr) function mode_field(full, oned) result(p)
      real(RealExt), intent(in), optional :: full(:, :, :), oned(:)
      integer :: j, jm, c
      j = new_buf(n_layer * n_mode)
      b(j)%v = 0.0_RealExt
      if (present(full) .or. present(oned)) then
        do jm = 1, n_mode
          do i = 1, n_layer
            ii = abs(layer_offset - i)
            c = i + n_layer * (jm - 1)
            do l = 1, n_profile
              if (present(full)) then
                if (last) then
                  b(j)%v(l, c) = full(ii, jm, list(l))
                else
                  b(j)%v(l, c) = full(list(l), ii, jm)
                end if
              else
                if (last) then
                  ll = n_mode * stride_aer * (list(l) - 1) + stride_aer * (jm - 1) + ii
                else
                  ll = n_profile * stride_aer * (jm - 1) + n_profile * (ii - 1) + list(l)
                end if
                b(j)%v(l, c) = oned(ll)
              end if
            end do
          end do
        end do
      end if
      p = c_loc(b(j)%v)
    end function mode_field

    ! a modal aerosol optical property per band (set_mode_opt_prop,
    ! socrates_set_aer.F90:494-566): buffer (profile, layer + n_layer*(mode-1)
    ! + n_layer*n_mode*(band-1))
    type(c_ptr) function mode_prop(full, oned) result(p)
      real(RealExt), intent(in), optional :: full(:, :, :, :), oned(:)
      integer :: j, jm, ib, c
      j = new_buf(n_layer * n_mode * int(nb))
      b(j)%v = 0.0_RealExt
      if (present(full) .or. present(oned)) then
        do ib = 1, nb
          do jm = 1, n_mode
            do i = 1, n_layer
              ii = abs(layer_offset - i)
              c = i + n_layer * (jm - 1) + n_layer * n_mode * (ib - 1)
              do l = 1, n_profile
                if (present(full)) then
                  if (last) then
                    b(j)%v(l, c) = full(ii, jm, ib, list(l))
                  else
                    b(j)%v(l, c) = full(list(l), ii, jm, ib)
                  end if
                else
                  if (last) then
                    ll = nb * n_mode * stride_aer * (list(l) - 1) + n_mode * stride_aer * (ib - 1) &
                       + stride_aer * (jm - 1) + ii
                  else
                    ll = n_mode * stride_aer * n_profile * (ib - 1) + stride_aer * n_profile * (jm - 1) &
                       + n_profile * (ii - 1) + list(l)
                  end if
                  b(j)%v(l, c) = oned(ll)
                end if
              end do
            end do
          end do
        end do
      end if
      p = c_loc(b(j)%v)
    end function mode_prop

    ! a per-profile field, taken at the listed profiles
    type(c_ptr) function col(x) result(p)
      real(RealExt), intent(in), optional :: x(:)
      integer :: j
      p = c_null_ptr
      if (.not. present(x)) return
      j = new_buf(1)
      b(j)%v(:, 1) = x(list)
      p = c_loc(b(j)%v)
    end function col

    logical function nonzero(full, oned)
      real(RealExt), intent(in), optional :: full(:, :), oned(:)
      nonzero = .false.
      if (present(full)) nonzero = any(full > 0.0_RealExt)
      if (present(oned)) nonzero = nonzero .or. any(oned > 0.0_RealExt)
    end function nonzero

    ! one gas: its form as set_gas_mix_ratio selects it
    ! (socrates_set_atm.F90:535-622): the full field, else the strided field,
    ! else the well-mixed ratio (also when l_well_mixed is set with one),
    ! else zero
    subroutine gas(ip, label, full, oned, mix, l_wm)
      integer, intent(in) :: ip
      character(len=*), intent(in) :: label
      real(RealExt), intent(in), optional :: full(:, :), oned(:), mix
      logical, intent(in), optional :: l_wm
      logical :: listed, well
      type(c_ptr) :: p
      integer :: s
      listed = .false.
      do s = 1, nabs
        listed = listed .or. types(s) == ip
      end do
      if (.not. listed) return
      well = .not. (present(full) .or. present(oned))
      if (present(mix) .and. present(l_wm)) well = well .or. l_wm
      p = c_null_ptr
      if (.not. well) p = lay(full, oned)
      select case (ip)
      case (1)
        a%h2o = p
        if (well .and. present(mix)) a%h2o_mix_ratio = mix
      case (2)
        a%co2 = p
        a%co2_mix_ratio = 0.0_RealExt
        if (well .and. present(mix)) a%co2_mix_ratio = mix
      case (3)
        a%o3 = p
        if (well .and. present(mix)) a%o3_mix_ratio = mix
      case default
        do s = 1, n_further
          if (further_ip(s) == ip) then
            a%gas_field(s) = p
            if (well .and. present(mix)) a%gas_mix_ratio(s) = mix
            return
          end if
        end do
        ! an absorber the library has no slot for: only a zero amount
        if (well) then
          if (present(mix)) then
            if (mix /= 0.0_RealExt) call fail('gas ' // label // ' is not on this path')
          end if
        else if (nonzero(full, oned)) then
          call fail('gas ' // label // ' is not on this path')
        end if
      end select
    end subroutine gas

    ! a dense output buffer of n2 rows per profile, handed to the library
    integer function out(p, n2)
      type(c_ptr), intent(out) :: p
      integer, intent(in) :: n2
      out = new_buf(n2)
      b(out)%v = 0.0_RealExt
      p = c_loc(b(out)%v)
    end function out

    ! a flux and its surface and top-of-atmosphere values, at the caller's
    ! indices of the associated arrays (sum_flux_channels,
    ! socrates_set_diag.F90:549-623)
    subroutine give_flux(x, surf, toa, j)
      real(RealExt), pointer :: x(:, :), surf(:), toa(:)
      integer, intent(in) :: j
      if (j == 0) return
      if (associated(x)) then
        do i = 0, n_layer
          ii = abs(level_offset - i)
          do l = 1, n_profile
            if (last) then
              x(ii, list(l)) = b(j)%v(l, i + 1)
            else
              x(list(l), ii) = b(j)%v(l, i + 1)
            end if
          end do
        end do
      end if
      if (associated(surf)) then
        do l = 1, n_profile
          surf(list(l)) = b(j)%v(l, n_layer + 1)
        end do
      end if
      if (associated(toa)) then
        do l = 1, n_profile
          toa(list(l)) = b(j)%v(l, 1)
        end do
      end if
    end subroutine give_flux

    ! heating rates at the caller's layers, zero in the array's layers
    ! outside 1..n_layer (socrates_set_diag.F90:139-225)
    subroutine give_heating(x, j)
      real(RealExt), pointer :: x(:, :)
      integer, intent(in) :: j
      do i = 1, n_layer
        ii = abs(layer_offset - i)
        do l = 1, n_profile
          if (last) then
            x(ii, list(l)) = b(j)%v(l, i)
          else
            x(list(l), ii) = b(j)%v(l, i)
          end if
        end do
      end do
      do l = 1, n_profile
        if (last) then
          do i = lbound(x, 1), 0
            x(i, list(l)) = 0.0_RealExt
          end do
          do i = n_layer + 1, ubound(x, 1)
            x(i, list(l)) = 0.0_RealExt
          end do
        else
          do i = lbound(x, 2), 0
            x(list(l), i) = 0.0_RealExt
          end do
          do i = n_layer + 1, ubound(x, 2)
            x(list(l), i) = 0.0_RealExt
          end do
        end if
      end do
    end subroutine give_heating

  end subroutine runes

end module socrates_runes
