! Fortran binding of the MI355X two-stream flux library (libsocgpu.so).
!
! Provides the module names, procedure names and keyword arguments of the
! reference interface for the two-stream flux path, so that a host model that
! calls SOCRATES through `runes` compiles against this file instead of the
! reference's interface_core:
!   realtype_rd            kinds RealK, RealExt          (realtype_rd.f90:14-20)
!   socrates_set_spectrum  set_spectrum, set_mcica       (socrates_set_spectrum.F90:29, 551)
!   socrates_runes         runes, StrDiag, ip_* codes    (socrates_runes.F90:47-683;
!                                                          socrates_def_diag.F90:31-76)
! Arguments of runes outside the path (other gases given as fields, aerosols,
! tiles, ice, l_invert/l_profile_last = .true., strided _1d forms, ...) are
! not declared: a call naming one does not compile, rather than being
! silently ignored. Well-mixed amounts of gases other than H2O, CO2 and O3 are
! accepted and must be absent from the spectral file (the reference reads
! only the gases its file lists, socrates_set_atm.F90:246-300).
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
  end type socgpu_runes_args

  ! SocgpuDiag
  type, bind(C) :: socgpu_diag
    type(c_ptr) :: flux_up = c_null_ptr, flux_down = c_null_ptr
    type(c_ptr) :: flux_direct = c_null_ptr, heating_rate = c_null_ptr
    integer(c_int) :: l_outputs_on_device = 0
  end type socgpu_diag

  interface
    integer(c_int) function socgpu_init(device) bind(C, name='socgpu_init')
      import :: c_int
      integer(c_int), value :: device
    end function socgpu_init
    integer(c_int) function socgpu_spectrum_register(name, path) &
        bind(C, name='socgpu_spectrum_register')
      import :: c_char, c_int
      character(kind=c_char), dimension(*) :: name, path
    end function socgpu_spectrum_register
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


module socrates_set_spectrum
  use, intrinsic :: iso_c_binding
  use socrates_gpu_c
  implicit none
  private
  public :: set_spectrum, set_mcica

contains

  ! Reads a spectral file and keeps it on the device under spectrum_name
  ! (socrates_set_spectrum.F90:29-149). n_instances and l_all_gases are
  ! accepted for compatibility; every gas of the file is kept.
  subroutine set_spectrum(n_instances, spectrum_name, spectral_file, l_all_gases)
    integer, intent(in), optional :: n_instances
    character(len=*), intent(in), optional :: spectrum_name
    character(len=*), intent(in), optional :: spectral_file
    logical, intent(in), optional :: l_all_gases
    if (.not. (present(spectrum_name) .and. present(spectral_file))) then
      if (present(n_instances)) return
      call check(-1_c_int, 'SET_SPECTRUM')
    end if
    call ensure_init()
    call check(socgpu_spectrum_register(c_str(spectrum_name), c_str(spectral_file)), &
               'SET_SPECTRUM')
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


module socrates_runes
  use, intrinsic :: iso_c_binding
  use realtype_rd, only: RealK, RealExt
  use socrates_gpu_c
  implicit none
  private
  public :: runes, StrDiag
  public :: ip_source_illuminate, ip_source_thermal, ip_solar, ip_infra_red
  public :: ip_cloud_off, ip_cloud_homogen, ip_cloud_ice_water, ip_cloud_conv_strat
  public :: ip_max_rand, ip_rand, ip_exponential_rand
  public :: ip_homogeneous, ip_mcica, ip_mcica_single_sampling, ip_mcica_optimal_sampling
  public :: ip_precision_fp64, ip_precision_mixed

  ! Option codes (rad_pcf.F90 values)
  integer, parameter :: ip_source_illuminate = 1, ip_source_thermal = 2
  integer, parameter :: ip_solar = 1, ip_infra_red = 2
  integer, parameter :: ip_cloud_off = 0, ip_cloud_homogen = 1
  integer, parameter :: ip_cloud_ice_water = 2, ip_cloud_conv_strat = 3
  integer, parameter :: ip_max_rand = 0, ip_rand = 1, ip_exponential_rand = 3
  integer, parameter :: ip_homogeneous = 0, ip_mcica = 2
  integer, parameter :: ip_mcica_single_sampling = 1, ip_mcica_optimal_sampling = 2
  ! Precision forms of the device solve (no reference counterpart)
  integer, parameter :: ip_precision_fp64 = 0, ip_precision_mixed = 1

  ! Caller-associated outputs (socrates_def_diag.F90:31-76, path subset)
  type :: StrDiag
    real(RealExt), pointer :: heating_rate(:, :) => null()
    real(RealExt), pointer :: flux_direct(:, :) => null()
    real(RealExt), pointer :: flux_down(:, :) => null()
    real(RealExt), pointer :: flux_up(:, :) => null()
  end type StrDiag

contains

  ! One radiation call on the device (socrates_runes.F90:47-1079 on the
  ! two-stream flux path). Array arguments have the reference's full-field
  ! shapes: (n_profile, n_layer), (n_profile, 0:n_layer), (n_profile),
  ! (n_profile, n_band); outputs through diag as in the reference.
  ! `precision` selects the device form (default FP64, the reference's).
  subroutine runes(n_profile, n_layer, diag, spectrum_name, n_subcol_gen, &
      p_layer, t_layer, p_level, t_level, mass, density, h2o, o3, &
      h2o_mix_ratio, co2_mix_ratio, o3_mix_ratio, n2o_mix_ratio, ch4_mix_ratio, &
      o2_mix_ratio, so2_mix_ratio, ocs_mix_ratio, cfc11_mix_ratio, cfc12_mix_ratio, &
      cfc113_mix_ratio, hcfc22_mix_ratio, hfc134a_mix_ratio, &
      t_ground, cos_zenith_angle, solar_irrad, &
      l_grey_albedo, grey_albedo, albedo_diff, albedo_dir, &
      cloud_frac, conv_frac, liq_mmr, liq_conv_mmr, liq_dim, liq_conv_dim, &
      cloud_vertical_decorr, conv_vertical_decorr, cloud_horizontal_rsd, &
      layer_heat_capacity, i_source, i_cloud_representation, i_overlap, i_inhom, &
      i_mcica_sampling, i_st_water, i_cnv_water, rand_seed, l_rayleigh, &
      l_invert, l_profile_last, precision)
    integer, intent(in) :: n_profile, n_layer
    type(StrDiag), intent(inout) :: diag
    character(len=*), intent(in), optional :: spectrum_name
    integer, intent(in), optional :: n_subcol_gen
    real(RealExt), intent(in), optional, target, contiguous :: p_layer(:, :), t_layer(:, :)
    real(RealExt), intent(in), optional, target, contiguous :: p_level(:, :), t_level(:, :)
    real(RealExt), intent(in), optional, target, contiguous :: mass(:, :), density(:, :)
    real(RealExt), intent(in), optional, target, contiguous :: h2o(:, :), o3(:, :)
    real(RealExt), intent(in), optional :: h2o_mix_ratio, co2_mix_ratio, o3_mix_ratio
    real(RealExt), intent(in), optional :: n2o_mix_ratio, ch4_mix_ratio, o2_mix_ratio
    real(RealExt), intent(in), optional :: so2_mix_ratio, ocs_mix_ratio, cfc11_mix_ratio
    real(RealExt), intent(in), optional :: cfc12_mix_ratio, cfc113_mix_ratio
    real(RealExt), intent(in), optional :: hcfc22_mix_ratio, hfc134a_mix_ratio
    real(RealExt), intent(in), optional, target, contiguous :: t_ground(:)
    real(RealExt), intent(in), optional, target, contiguous :: cos_zenith_angle(:)
    real(RealExt), intent(in), optional, target, contiguous :: solar_irrad(:)
    logical, intent(in), optional :: l_grey_albedo
    real(RealExt), intent(in), optional :: grey_albedo
    real(RealExt), intent(in), optional, target, contiguous :: albedo_diff(:, :), albedo_dir(:, :)
    real(RealExt), intent(in), optional, target, contiguous :: cloud_frac(:, :), conv_frac(:, :)
    real(RealExt), intent(in), optional, target, contiguous :: liq_mmr(:, :), liq_conv_mmr(:, :)
    real(RealExt), intent(in), optional, target, contiguous :: liq_dim(:, :), liq_conv_dim(:, :)
    real(RealExt), intent(in), optional :: cloud_vertical_decorr, conv_vertical_decorr
    real(RealExt), intent(in), optional :: cloud_horizontal_rsd
    real(RealExt), intent(in), optional, target, contiguous :: layer_heat_capacity(:, :)
    integer, intent(in), optional :: i_source, i_cloud_representation, i_overlap, i_inhom
    integer, intent(in), optional :: i_mcica_sampling, i_st_water, i_cnv_water
    integer, intent(in), optional, target, contiguous :: rand_seed(:)
    logical, intent(in), optional :: l_rayleigh, l_invert, l_profile_last
    integer, intent(in), optional :: precision

    type(socgpu_runes_args) :: a
    type(socgpu_diag) :: d
    ! l_invert: bottom-up caller profiles, flipped into top-down copies and
    ! the outputs flipped back (socrates_set_atm.F90:218-225,
    ! socrates_set_diag.F90:113-151)
    type :: flip_buf
      real(RealExt), allocatable :: v(:, :)
    end type flip_buf
    type(flip_buf), target :: fin(20), fout(4)
    integer :: nfin
    logical :: inv, last
    integer(c_int) :: prec, nb, nabs, types(32), status
    integer(c_int), allocatable, target :: seeds(:)
    character(len=64) :: name

    call ensure_init()
    inv = .false.
    if (present(l_invert)) inv = l_invert
    nfin = 0
    ! l_profile_last: fields (levels or layers or bands, n_profile), handed
    ! to the library transposed and the outputs transposed back
    ! (socrates_set_atm.F90:449-470, socrates_set_bound.F90:156-172)
    last = .false.
    if (present(l_profile_last)) last = l_profile_last
    if (.not. present(spectrum_name)) call fail('spectrum_name is required')
    name = spectrum_name
    call check(socgpu_spectrum_query(c_str(name), nb, nabs, types, 32_c_int), 'RUNES')
    if (present(h2o_mix_ratio) .and. .not. present(h2o)) call fail('h2o_mix_ratio: give h2o')
    if (present(o3_mix_ratio) .and. .not. present(o3)) call fail('o3_mix_ratio: give o3')

    a%n_profile = n_profile
    a%n_layer = n_layer
    if (present(i_source)) a%i_source = i_source
    if (present(l_rayleigh)) a%l_rayleigh = merge(1, 0, l_rayleigh)
    if (present(i_cloud_representation)) a%i_cloud_representation = i_cloud_representation
    if (present(i_overlap)) a%i_overlap = i_overlap
    if (present(i_inhom)) a%i_inhom = i_inhom
    if (present(i_mcica_sampling)) a%i_mcica_sampling = i_mcica_sampling
    if (present(i_st_water)) a%i_st_water = i_st_water
    if (present(i_cnv_water)) a%i_cnv_water = i_cnv_water
    if (present(n_subcol_gen)) a%n_subcol_gen = n_subcol_gen
    if (present(co2_mix_ratio)) a%co2_mix_ratio = co2_mix_ratio
    ! further well-mixed gases (absent = 0, socrates_set_atm.F90:596-617)
    if (present(n2o_mix_ratio)) a%gas_mix_ratio(1) = n2o_mix_ratio
    if (present(ch4_mix_ratio)) a%gas_mix_ratio(2) = ch4_mix_ratio
    if (present(o2_mix_ratio)) a%gas_mix_ratio(3) = o2_mix_ratio
    if (present(so2_mix_ratio)) a%gas_mix_ratio(4) = so2_mix_ratio
    if (present(ocs_mix_ratio)) a%gas_mix_ratio(5) = ocs_mix_ratio
    if (present(cfc11_mix_ratio)) a%gas_mix_ratio(6) = cfc11_mix_ratio
    if (present(cfc12_mix_ratio)) a%gas_mix_ratio(7) = cfc12_mix_ratio
    if (present(cfc113_mix_ratio)) a%gas_mix_ratio(8) = cfc113_mix_ratio
    if (present(hcfc22_mix_ratio)) a%gas_mix_ratio(9) = hcfc22_mix_ratio
    if (present(hfc134a_mix_ratio)) a%gas_mix_ratio(10) = hfc134a_mix_ratio
    if (present(cloud_vertical_decorr)) a%cloud_vertical_decorr = cloud_vertical_decorr
    if (present(conv_vertical_decorr)) a%conv_vertical_decorr = conv_vertical_decorr
    if (present(cloud_horizontal_rsd)) a%cloud_horizontal_rsd = cloud_horizontal_rsd
    if (present(l_grey_albedo)) a%l_grey_albedo = merge(1, 0, l_grey_albedo)
    if (present(grey_albedo)) then
      a%grey_albedo = grey_albedo
    else
      a%l_grey_albedo = 0
    end if

    call field(p_layer, n_layer, a%p_layer, 'p_layer', .true.)
    call field(t_layer, n_layer, a%t_layer, 't_layer', .true.)
    call field(mass, n_layer, a%mass, 'mass', .true.)
    call field(density, n_layer, a%density, 'density', .true.)
    call field(layer_heat_capacity, n_layer, a%layer_heat_capacity, 'layer_heat_capacity', .true.)
    call field(h2o, n_layer, a%h2o, 'h2o', .true.)
    call field(o3, n_layer, a%o3, 'o3', .true.)
    call field(p_level, n_layer + 1, a%p_level, 'p_level', .true.)
    call field(t_level, n_layer + 1, a%t_level, 't_level', .true.)
    call field(albedo_diff, nb, a%albedo_diff, 'albedo_diff', .false.)
    call field(albedo_dir, nb, a%albedo_dir, 'albedo_dir', .false.)
    call field(cloud_frac, n_layer, a%cloud_frac, 'cloud_frac', .true.)
    call field(liq_mmr, n_layer, a%liq_mmr, 'liq_mmr', .true.)
    call field(liq_dim, n_layer, a%liq_dim, 'liq_dim', .true.)
    call field(conv_frac, n_layer, a%conv_frac, 'conv_frac', .true.)
    call field(liq_conv_mmr, n_layer, a%liq_conv_mmr, 'liq_conv_mmr', .true.)
    call field(liq_conv_dim, n_layer, a%liq_conv_dim, 'liq_conv_dim', .true.)
    call column(t_ground, a%t_ground, 't_ground')
    call column(cos_zenith_angle, a%cos_zenith_angle, 'cos_zenith_angle')
    call column(solar_irrad, a%solar_irrad, 'solar_irrad')
    if (present(rand_seed)) then
      if (size(rand_seed) /= n_profile) call fail('rand_seed must have n_profile elements')
      allocate(seeds(n_profile))
      seeds = int(rand_seed, c_int)
      a%rand_seed = c_loc(seeds)
    end if

    call output(diag%flux_up, n_layer + 1, d%flux_up, 'flux_up', 1)
    call output(diag%flux_down, n_layer + 1, d%flux_down, 'flux_down', 2)
    call output(diag%flux_direct, n_layer + 1, d%flux_direct, 'flux_direct', 3)
    call output(diag%heating_rate, n_layer, d%heating_rate, 'heating_rate', 4)

    prec = ip_precision_fp64
    if (present(precision)) prec = precision
    status = socgpu_runes(c_str(name), a, d, prec)
    call check(status, 'RUNES')
    if (associated(diag%flux_up)) call give_back(diag%flux_up, 1, n_layer + 1)
    if (associated(diag%flux_down)) call give_back(diag%flux_down, 2, n_layer + 1)
    if (associated(diag%flux_direct)) call give_back(diag%flux_direct, 3, n_layer + 1)
    if (associated(diag%heating_rate)) call give_back(diag%heating_rate, 4, n_layer)

  contains

    subroutine fail(msg)
      character(len=*), intent(in) :: msg
      write(*, '(a, a)') '? Error from routine: RUNES: ', msg
      error stop 1
    end subroutine fail

    subroutine field(x, n2, p, label, vertical)
      real(RealExt), intent(in), optional, target, contiguous :: x(:, :)
      integer, intent(in) :: n2
      type(c_ptr), intent(out) :: p
      character(len=*), intent(in) :: label
      logical, intent(in) :: vertical
      p = c_null_ptr
      if (.not. present(x)) return
      if (last) then
        if (size(x, 1) /= n2 .or. size(x, 2) /= n_profile) &
          call fail(label // ': shape must be (levels or layers or bands, n_profile)')
      else if (size(x, 1) /= n_profile .or. size(x, 2) /= n2) then
        call fail(label // ': shape must be (n_profile, levels or layers or bands)')
      end if
      if (last .or. (inv .and. vertical)) then
        nfin = nfin + 1
        allocate(fin(nfin)%v(n_profile, n2))
        if (last) then
          fin(nfin)%v = transpose(x)
        else
          fin(nfin)%v = x
        end if
        if (inv .and. vertical) fin(nfin)%v = fin(nfin)%v(:, n2:1:-1)
        p = c_loc(fin(nfin)%v)
      else
        p = c_loc(x)
      end if
    end subroutine field

    ! an output written to a library buffer, flipped and / or transposed
    ! into the caller's array
    subroutine give_back(x, k, n2)
      real(RealExt), pointer :: x(:, :)
      integer, intent(in) :: k, n2
      if (.not. (inv .or. last)) return
      if (inv) fout(k)%v = fout(k)%v(:, n2:1:-1)
      if (last) then
        x = transpose(fout(k)%v)
      else
        x = fout(k)%v
      end if
    end subroutine give_back

    subroutine column(x, p, label)
      real(RealExt), intent(in), optional, target, contiguous :: x(:)
      type(c_ptr), intent(out) :: p
      character(len=*), intent(in) :: label
      p = c_null_ptr
      if (.not. present(x)) return
      if (size(x) /= n_profile) call fail(label // ': size must be n_profile')
      p = c_loc(x)
    end subroutine column

    subroutine output(x, n2, p, label, k)
      real(RealExt), pointer :: x(:, :)
      integer, intent(in) :: n2, k
      type(c_ptr), intent(out) :: p
      character(len=*), intent(in) :: label
      p = c_null_ptr
      if (.not. associated(x)) return
      if (last) then
        if (size(x, 1) /= n2 .or. size(x, 2) /= n_profile) &
          call fail(label // ': shape must be (0:n_layer, n_profile) or (n_layer, n_profile)')
      else if (size(x, 1) /= n_profile .or. size(x, 2) /= n2) then
        call fail(label // ': shape must be (n_profile, 0:n_layer) or (n_profile, n_layer)')
      end if
      if (.not. is_contiguous(x)) call fail(label // ' must be contiguous')
      if (inv .or. last) then
        allocate(fout(k)%v(n_profile, n2))
        p = c_loc(fout(k)%v)
      else
        p = c_loc(x)
      end if
    end subroutine output

  end subroutine runes

end module socrates_runes
