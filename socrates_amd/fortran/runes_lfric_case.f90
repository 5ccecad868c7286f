! Example host program in the calling form of a host model's radiation
! driver (the reference's runes_nc.f90:2133-2290): spectra registered with
! per-gas flags, and runes called on a profile list with strided
! one-dimensional fields stored bottom-up with profiles last (l_invert,
! l_profile_last, n_layer_stride), every gas given as a strided field or a
! well-mixed ratio with its l_*_well_mixed switch, and clear-sky, surface and
! top-of-atmosphere diagnostics. The same source builds against this
! library's Fortran binding and against the reference, so that both read the
! same caller arrays.
!
! The call's n_profile columns of a case file (socrates_amd/case_io.py) are
! placed at positions n_full - l (l = 1..n_profile) of n_full = n_profile + 3
! stored columns, whose unused entries hold values no correct indexing
! reads. Outputs are written for the call's columns in the case file's order
! (top-down, profile fastest): flux_up, flux_down, flux_direct,
! heating_rate, flux_up_clear, flux_down_clear, flux_direct_clear, then the
! six surface and top-of-atmosphere values (up, down, direct at the
! surface, then at the top).
!
! Usage: runes_lfric_case <sw_spectral_file> <lw_spectral_file> <mcica_file|-> \
!                         <case.bin> <out.bin> [drop_gas_identifier]
! With a gas identifier (rad_pcf.F90 ip_*), that absorber is left out of
! the spectra by set_spectrum's gas flags.
program runes_lfric_case

  use socrates_set_spectrum, only: set_spectrum, set_mcica
  use socrates_runes, only: runes, StrDiag
  use def_spectrum, only: StrSpecData
  use realtype_rd, only: RealExt

  implicit none

  character(len=1024) :: sw_file, lw_file, mcica_file, case_file, out_file, arg
  integer :: iu, n_profile, n_layer, i_source, l_cloud, i_overlap, l_rayleigh
  integer :: magic, version, n_band, i_st_water, i_st_ice, i_cloud_repr
  integer :: i_inhom, i_mcica_sampling, n_subcol_gen, i_solver_kind
  integer :: lfg, lco2, n_full, l, i, ii, ib, pos, drop
  real(RealExt) :: co2_mix_ratio, cloud_vertical_decorr, cloud_horizontal_rsd
  real(RealExt) :: conv_vertical_decorr, wm(10)
  real(RealExt), allocatable, dimension(:,:) :: p_layer, t_layer, mass, density, &
    layer_heat_capacity, h2o, o3, p_level, t_level, albedo_diff, albedo_dir, &
    cloud_frac, liq_mmr, liq_dim, conv_frac, liq_conv_mmr, liq_conv_dim, &
    flux_ground, co2
  real(RealExt), allocatable, dimension(:) :: t_ground, cos_zen, solar_irrad
  integer, allocatable :: rand_seed(:), seed_full(:), profile_list(:)
  ! stored (host-model) forms
  real(RealExt), allocatable, dimension(:) :: p_layer_1d, t_layer_1d, mass_1d, &
    density_1d, hc_1d, h2o_1d, o3_1d, co2_1d, ch4_1d, n2o_1d, p_level_1d, &
    t_level_1d, albedo_diff_1d, albedo_dir_1d, flux_ground_1d, cloud_frac_1d, &
    liq_mmr_1d, liq_dim_1d, conv_frac_1d, liq_conv_mmr_1d, liq_conv_dim_1d, &
    ice_mmr_1d, t_ground_f, cos_zen_f, solar_irrad_f
  real(RealExt), allocatable, target, dimension(:,:) :: flux_up, flux_down, &
    flux_direct, heating_rate, flux_up_clear, flux_down_clear, flux_direct_clear
  real(RealExt), allocatable, target, dimension(:) :: up_surf, down_surf, &
    direct_surf, up_toa, down_toa, direct_toa
  real(RealExt), allocatable :: w(:,:)
  real(RealExt), parameter :: junk = -7.5e3_RealExt
  type(StrDiag) :: diag
  type(StrSpecData) :: sw_spec, lw_spec
  logical :: keep(67)
  character(len=2) :: name
  ! fields of case-file versions 5-7 this host does not pass on, and the
  ! modal aerosols of version 8 (stored as the host model stores them)
  integer :: lwf(10), lice, iscat, n_aer, jm
  real(RealExt), allocatable :: skip(:, :), aer_mix(:, :, :), aer_abs(:, :, :, :), &
    aer_sca(:, :, :, :), aer_asy(:, :, :, :)
  real(RealExt), allocatable, dimension(:) :: aer_mix_1d, aer_abs_1d, aer_sca_1d, aer_asy_1d

  call get_command_argument(1, sw_file)
  call get_command_argument(2, lw_file)
  call get_command_argument(3, mcica_file)
  call get_command_argument(4, case_file)
  call get_command_argument(5, out_file)
  drop = 0
  if (command_argument_count() >= 6) then
    call get_command_argument(6, arg)
    read(arg, *) drop
  end if

  open(newunit=iu, file=trim(case_file), access='stream', form='unformatted', &
       status='old')
  read(iu) magic, version, n_profile, n_layer, n_band, i_source, l_cloud, &
    i_cloud_repr, i_overlap, i_inhom, i_mcica_sampling, l_rayleigh, &
    i_st_water, i_st_ice, n_subcol_gen, i_solver_kind
  if (magic /= 1397703491) stop 'bad case file magic'
  allocate(p_layer(n_profile, n_layer), t_layer(n_profile, n_layer), &
    mass(n_profile, n_layer), density(n_profile, n_layer), &
    layer_heat_capacity(n_profile, n_layer), h2o(n_profile, n_layer), &
    o3(n_profile, n_layer), p_level(n_profile, 0:n_layer), &
    t_level(n_profile, 0:n_layer), albedo_diff(n_profile, n_band), &
    albedo_dir(n_profile, n_band), cloud_frac(n_profile, n_layer), &
    liq_mmr(n_profile, n_layer), liq_dim(n_profile, n_layer), &
    conv_frac(n_profile, n_layer), liq_conv_mmr(n_profile, n_layer), &
    liq_conv_dim(n_profile, n_layer), flux_ground(n_profile, n_band), &
    co2(n_profile, n_layer), &
    t_ground(n_profile), cos_zen(n_profile), solar_irrad(n_profile), &
    rand_seed(n_profile))
  read(iu) co2_mix_ratio, cloud_vertical_decorr, cloud_horizontal_rsd, &
    conv_vertical_decorr
  read(iu) p_layer, t_layer, mass, density, layer_heat_capacity, h2o, o3
  read(iu) p_level, t_level
  read(iu) t_ground, cos_zen, solar_irrad
  read(iu) albedo_diff, albedo_dir
  read(iu) cloud_frac, liq_mmr, liq_dim
  read(iu) conv_frac, liq_conv_mmr, liq_conv_dim
  read(iu) rand_seed
  wm = 0.0_RealExt
  if (version >= 2) read(iu) wm
  lfg = 0
  lco2 = 0
  flux_ground = 0.0_RealExt
  if (version >= 3) read(iu) lfg, flux_ground
  if (version >= 4) then
    read(iu) lco2
    if (lco2 /= 0) read(iu) co2
  end if
  allocate(skip(n_profile, n_layer))
  if (version >= 5) then
    read(iu) lwf
    do i = 1, 10
      if (lwf(i) /= 0) read(iu) skip
    end do
  end if
  if (version >= 6) then
    read(iu) lice
    if (lice /= 0) read(iu) skip, skip, skip, skip
  end if
  if (version >= 7) read(iu) iscat
  n_aer = 0
  if (version >= 8) then
    read(iu) n_aer
    if (n_aer > 0) then
      allocate(aer_mix(n_profile, n_layer, n_aer), aer_abs(n_profile, n_layer, n_aer, n_band), &
        aer_sca(n_profile, n_layer, n_aer, n_band), aer_asy(n_profile, n_layer, n_aer, n_band))
      read(iu) aer_mix, aer_abs, aer_sca, aer_asy
    end if
  end if
  close(iu)

  ! host-model storage: n_full columns, the call's column l at pos(l)
  n_full = n_profile + 3
  allocate(profile_list(n_profile))
  do l = 1, n_profile
    profile_list(l) = n_full - l
  end do
  p_layer_1d = lay1d(p_layer); t_layer_1d = lay1d(t_layer)
  mass_1d = lay1d(mass); density_1d = lay1d(density)
  hc_1d = lay1d(layer_heat_capacity)
  h2o_1d = lay1d(h2o); o3_1d = lay1d(o3)
  co2_1d = lay1d(co2)
  p_level_1d = lev1d(p_level); t_level_1d = lev1d(t_level)
  albedo_diff_1d = band1d(albedo_diff); albedo_dir_1d = band1d(albedo_dir)
  flux_ground_1d = band1d(flux_ground)
  allocate(w(n_profile, n_layer))
  ! CH4 as a field of its well-mixed value; N2O's field is junk, switched
  ! off by its l_n2o_well_mixed
  w = wm(2)
  ch4_1d = lay1d(w)
  w = junk
  n2o_1d = lay1d(w)
  cloud_frac_1d = lay1d(cloud_frac); liq_mmr_1d = lay1d(liq_mmr)
  liq_dim_1d = lay1d(liq_dim); conv_frac_1d = lay1d(conv_frac)
  liq_conv_mmr_1d = lay1d(liq_conv_mmr); liq_conv_dim_1d = lay1d(liq_conv_dim)
  w = 0.0_RealExt
  ice_mmr_1d = lay1d(w)
  if (n_aer > 0) then
    aer_mix_1d = mode1d(aer_mix)
    aer_abs_1d = prop1d(aer_abs); aer_sca_1d = prop1d(aer_sca); aer_asy_1d = prop1d(aer_asy)
  end if
  allocate(t_ground_f(n_full), cos_zen_f(n_full), solar_irrad_f(n_full), seed_full(n_full))
  t_ground_f = junk; cos_zen_f = 0.5_RealExt; solar_irrad_f = junk; seed_full = -1
  do l = 1, n_profile
    pos = profile_list(l)
    t_ground_f(pos) = t_ground(l)
    cos_zen_f(pos) = cos_zen(l)
    solar_irrad_f(pos) = solar_irrad(l)
    seed_full(pos) = rand_seed(l)
  end do

  ! every absorber of the files kept, but the one dropped
  keep = .true.
  if (drop > 0) keep(drop) = .false.
  call set_spectrum(spectrum_name='sw', spectral_file=trim(sw_file), &
    l_h2o=keep(1), l_co2=keep(2), l_o3=keep(3), l_n2o=keep(4), l_ch4=keep(6), &
    l_o2=keep(7), l_so2=keep(9), l_cfc11=keep(14), l_cfc12=keep(15), &
    l_cfc113=keep(16), l_hcfc22=keep(17), l_hfc134a=keep(19), l_ocs=keep(25))
  call set_spectrum(spectrum_name='lw', spectral_file=trim(lw_file), &
    l_h2o=keep(1), l_co2=keep(2), l_o3=keep(3), l_n2o=keep(4), l_ch4=keep(6), &
    l_o2=keep(7), l_so2=keep(9), l_cfc11=keep(14), l_cfc12=keep(15), &
    l_cfc113=keep(16), l_hcfc22=keep(17), l_hfc134a=keep(19), l_ocs=keep(25))
  ! and a spectrum handle, used for the cloud-free calls
  call set_spectrum(spectrum=sw_spec, spectral_file=trim(sw_file), &
    l_h2o=keep(1), l_co2=keep(2), l_o3=keep(3), l_n2o=keep(4), l_ch4=keep(6), &
    l_o2=keep(7), l_so2=keep(9), l_cfc11=keep(14), l_cfc12=keep(15), &
    l_cfc113=keep(16), l_hcfc22=keep(17), l_hfc134a=keep(19), l_ocs=keep(25))
  call set_spectrum(spectrum=lw_spec, spectral_file=trim(lw_file), &
    l_h2o=keep(1), l_co2=keep(2), l_o3=keep(3), l_n2o=keep(4), l_ch4=keep(6), &
    l_o2=keep(7), l_so2=keep(9), l_cfc11=keep(14), l_cfc12=keep(15), &
    l_cfc113=keep(16), l_hcfc22=keep(17), l_hfc134a=keep(19), l_ocs=keep(25))
  if (trim(mcica_file) /= '-') then
    call set_mcica(trim(mcica_file), sw_spectrum_name='sw', lw_spectrum_name='lw')
  end if

  allocate(flux_up(0:n_layer, n_full), flux_down(0:n_layer, n_full), &
    flux_direct(0:n_layer, n_full), heating_rate(0:n_layer + 1, n_full), &
    flux_up_clear(0:n_layer, n_full), flux_down_clear(0:n_layer, n_full), &
    flux_direct_clear(0:n_layer, n_full), up_surf(n_full), down_surf(n_full), &
    direct_surf(n_full), up_toa(n_full), down_toa(n_full), direct_toa(n_full))
  flux_up = junk; flux_down = junk; flux_direct = 0.0_RealExt
  heating_rate = junk; flux_up_clear = junk; flux_down_clear = junk
  flux_direct_clear = 0.0_RealExt
  up_surf = junk; down_surf = junk; direct_surf = 0.0_RealExt
  up_toa = junk; down_toa = junk; direct_toa = 0.0_RealExt
  diag%flux_up => flux_up
  diag%flux_down => flux_down
  diag%heating_rate => heating_rate
  diag%flux_up_clear => flux_up_clear
  diag%flux_down_clear => flux_down_clear
  diag%flux_up_surf => up_surf
  diag%flux_down_surf => down_surf
  diag%flux_up_toa => up_toa
  diag%flux_down_toa => down_toa
  if (i_source == 1) then
    diag%flux_direct => flux_direct
    diag%flux_direct_clear => flux_direct_clear
    diag%flux_direct_surf => direct_surf
    diag%flux_direct_toa => direct_toa
    name = 'sw'
  else
    name = 'lw'
  end if

  if (l_cloud == 0) then
    if (i_source == 1) then
      call clear_call(sw_spec)
    else
      call clear_call(lw_spec)
    end if
  else if (i_solver_kind == 0) then
    call runes(n_profile, n_layer, diag, spectrum_name=name, &
      i_source=i_source, profile_list=profile_list, &
      n_layer_stride=n_layer + 1, l_invert=.true., l_profile_last=.true., &
      n_cloud_layer=n_layer, &
      p_layer_1d=p_layer_1d, t_layer_1d=t_layer_1d, p_level_1d=p_level_1d, &
      t_level_1d=t_level_1d, mass_1d=mass_1d, density_1d=density_1d, &
      h2o_1d=h2o_1d, o3_1d=o3_1d, co2_1d=co2_1d, ch4_1d=ch4_1d, n2o_1d=n2o_1d, &
      h2o_mix_ratio=0.0_RealExt, l_h2o_well_mixed=.false., &
      co2_mix_ratio=co2_mix_ratio, l_co2_well_mixed=(lco2 == 0), &
      n2o_mix_ratio=wm(1), l_n2o_well_mixed=.true., &
      ch4_mix_ratio=junk, l_ch4_well_mixed=.false., &
      o2_mix_ratio=wm(3), so2_mix_ratio=wm(4), ocs_mix_ratio=wm(5), &
      cfc11_mix_ratio=wm(6), cfc12_mix_ratio=wm(7), cfc113_mix_ratio=wm(8), &
      hcfc22_mix_ratio=wm(9), hfc134a_mix_ratio=wm(10), &
      l_flux_ground=(lfg /= 0), t_ground=t_ground_f, flux_ground_1d=flux_ground_1d, &
      cos_zenith_angle=cos_zen_f, solar_irrad=solar_irrad_f, &
      albedo_diff_1d=albedo_diff_1d, albedo_dir_1d=albedo_dir_1d, &
      l_rayleigh=(l_rayleigh /= 0), layer_heat_capacity_1d=hc_1d, &
      cloud_frac_1d=cloud_frac_1d, liq_mmr_1d=liq_mmr_1d, liq_dim_1d=liq_dim_1d, &
      ice_mmr_1d=ice_mmr_1d, ice_dim_1d=liq_dim_1d, &
      cloud_vertical_decorr=cloud_vertical_decorr, &
      cloud_horizontal_rsd=cloud_horizontal_rsd, &
      i_cloud_representation=i_cloud_repr, i_overlap=i_overlap, &
      i_inhom=i_inhom, i_mcica_sampling=i_mcica_sampling, &
      i_st_water=i_st_water, i_st_ice=i_st_ice, &
      n_subcol_gen=n_subcol_gen, rand_seed=seed_full, &
      l_aerosol_mode=(n_aer > 0), n_aer_mode=n_aer, aer_mix_ratio_1d=aer_mix_1d, &
      aer_absorption_1d=aer_abs_1d, aer_scattering_1d=aer_sca_1d, &
      aer_asymmetry_1d=aer_asy_1d)
  else
    call runes(n_profile, n_layer, diag, spectrum_name=name, &
      i_source=i_source, profile_list=profile_list, &
      n_layer_stride=n_layer + 1, l_invert=.true., l_profile_last=.true., &
      p_layer_1d=p_layer_1d, t_layer_1d=t_layer_1d, p_level_1d=p_level_1d, &
      t_level_1d=t_level_1d, mass_1d=mass_1d, density_1d=density_1d, &
      h2o_1d=h2o_1d, o3_1d=o3_1d, co2_1d=co2_1d, ch4_1d=ch4_1d, n2o_1d=n2o_1d, &
      co2_mix_ratio=co2_mix_ratio, l_co2_well_mixed=(lco2 == 0), &
      n2o_mix_ratio=wm(1), l_n2o_well_mixed=.true., &
      o2_mix_ratio=wm(3), so2_mix_ratio=wm(4), ocs_mix_ratio=wm(5), &
      cfc11_mix_ratio=wm(6), cfc12_mix_ratio=wm(7), cfc113_mix_ratio=wm(8), &
      hcfc22_mix_ratio=wm(9), hfc134a_mix_ratio=wm(10), &
      l_flux_ground=(lfg /= 0), t_ground=t_ground_f, flux_ground_1d=flux_ground_1d, &
      cos_zenith_angle=cos_zen_f, solar_irrad=solar_irrad_f, &
      albedo_diff_1d=albedo_diff_1d, albedo_dir_1d=albedo_dir_1d, &
      l_rayleigh=(l_rayleigh /= 0), layer_heat_capacity_1d=hc_1d, &
      cloud_frac_1d=cloud_frac_1d, conv_frac_1d=conv_frac_1d, &
      liq_mmr_1d=liq_mmr_1d, liq_dim_1d=liq_dim_1d, &
      liq_conv_mmr_1d=liq_conv_mmr_1d, liq_conv_dim_1d=liq_conv_dim_1d, &
      ice_mmr_1d=ice_mmr_1d, ice_dim_1d=liq_dim_1d, &
      ice_conv_mmr_1d=ice_mmr_1d, ice_conv_dim_1d=liq_dim_1d, &
      cloud_vertical_decorr=cloud_vertical_decorr, &
      conv_vertical_decorr=conv_vertical_decorr, &
      i_cloud_representation=i_cloud_repr, i_overlap=i_overlap, &
      i_st_water=i_st_water, i_st_ice=i_st_ice, &
      i_cnv_water=i_st_water, i_cnv_ice=i_st_ice, &
      l_aerosol_mode=(n_aer > 0), n_aer_mode=n_aer, aer_mix_ratio_1d=aer_mix_1d, &
      aer_absorption_1d=aer_abs_1d, aer_scattering_1d=aer_sca_1d, &
      aer_asymmetry_1d=aer_asy_1d)
  end if

  ! back to the case file's order: top-down, the call's columns
  open(newunit=iu, file=trim(out_file), access='stream', form='unformatted', &
       status='replace')
  call put_levels(flux_up)
  call put_levels(flux_down)
  call put_levels(flux_direct)
  do i = 1, n_layer
    ii = n_layer + 1 - i
    write(iu) (heating_rate(ii, profile_list(l)), l = 1, n_profile)
  end do
  call put_levels(flux_up_clear)
  call put_levels(flux_down_clear)
  call put_levels(flux_direct_clear)
  write(iu) up_surf(profile_list), down_surf(profile_list), direct_surf(profile_list)
  write(iu) up_toa(profile_list), down_toa(profile_list), direct_toa(profile_list)
  ! the heating rate's further layers and the unused columns
  write(iu) heating_rate(0, profile_list), heating_rate(n_layer + 1, profile_list)
  write(iu) flux_up(:, 1), flux_up(:, 2), flux_up(:, n_full)
  close(iu)

contains

  subroutine clear_call(spec)
    type(StrSpecData), intent(in) :: spec
    call runes(n_profile, n_layer, diag, spec, &
      i_source=i_source, profile_list=profile_list, &
      n_layer_stride=n_layer + 1, l_invert=.true., l_profile_last=.true., &
      p_layer_1d=p_layer_1d, t_layer_1d=t_layer_1d, p_level_1d=p_level_1d, &
      t_level_1d=t_level_1d, mass_1d=mass_1d, density_1d=density_1d, &
      h2o_1d=h2o_1d, o3_1d=o3_1d, co2_1d=co2_1d, ch4_1d=ch4_1d, n2o_1d=n2o_1d, &
      co2_mix_ratio=co2_mix_ratio, l_co2_well_mixed=(lco2 == 0), &
      n2o_mix_ratio=wm(1), l_n2o_well_mixed=.true., &
      o2_mix_ratio=wm(3), so2_mix_ratio=wm(4), ocs_mix_ratio=wm(5), &
      cfc11_mix_ratio=wm(6), cfc12_mix_ratio=wm(7), cfc113_mix_ratio=wm(8), &
      hcfc22_mix_ratio=wm(9), hfc134a_mix_ratio=wm(10), &
      l_flux_ground=(lfg /= 0), t_ground=t_ground_f, flux_ground_1d=flux_ground_1d, &
      cos_zenith_angle=cos_zen_f, solar_irrad=solar_irrad_f, &
      albedo_diff_1d=albedo_diff_1d, albedo_dir_1d=albedo_dir_1d, &
      l_rayleigh=(l_rayleigh /= 0), layer_heat_capacity_1d=hc_1d, &
      l_aerosol_mode=(n_aer > 0), n_aer_mode=n_aer, aer_mix_ratio_1d=aer_mix_1d, &
      aer_absorption_1d=aer_abs_1d, aer_scattering_1d=aer_sca_1d, &
      aer_asymmetry_1d=aer_asy_1d)
  end subroutine clear_call

  ! a layer field stored bottom-up per column, n_layer + 1 entries a column
  ! (entry 0 unused), profiles last
  function lay1d(x) result(y)
    real(RealExt), intent(in) :: x(:, :)
    real(RealExt), allocatable :: y(:)
    integer :: ll
    allocate(y((n_layer + 1) * n_full))
    y = junk
    do l = 1, n_profile
      do i = 1, n_layer
        ll = (n_layer + 1) * (profile_list(l) - 1) + (n_layer + 1 - i)
        y(ll) = x(l, i)
      end do
    end do
  end function lay1d

  ! a level field stored bottom-up, n_layer + 1 entries a column
  function lev1d(x) result(y)
    real(RealExt), intent(in) :: x(:, 0:)
    real(RealExt), allocatable :: y(:)
    integer :: ll
    allocate(y((n_layer + 1) * n_full))
    y = junk
    do l = 1, n_profile
      do i = 0, n_layer
        ll = (n_layer + 1) * (profile_list(l) - 1) + (n_layer - i) + 1
        y(ll) = x(l, i)
      end do
    end do
  end function lev1d

  ! a per-band field, bands fastest
  function band1d(x) result(y)
    real(RealExt), intent(in) :: x(:, :)
    real(RealExt), allocatable :: y(:)
    allocate(y(n_band * n_full))
    y = junk
    do l = 1, n_profile
      do ib = 1, n_band
        y(n_band * (profile_list(l) - 1) + ib) = x(l, ib)
      end do
    end do
  end function band1d

  ! a modal aerosol mixing ratio stored as the host model stores it
  ! (socrates_set_aer.F90:466-473 with l_profile_last and l_invert):
  ! per column, per mode, n_layer + 1 entries bottom-up
  function mode1d(x) result(y)
    real(RealExt), intent(in) :: x(:, :, :)
    real(RealExt), allocatable :: y(:)
    integer :: ll
    allocate(y(n_aer * (n_layer + 1) * n_full))
    y = junk
    do l = 1, n_profile
      do jm = 1, n_aer
        do i = 1, n_layer
          ll = n_aer * (n_layer + 1) * (profile_list(l) - 1) + (n_layer + 1) * (jm - 1) &
             + (n_layer + 1 - i)
          y(ll) = x(l, i, jm)
        end do
      end do
    end do
  end function mode1d

  ! a modal aerosol optical property (socrates_set_aer.F90:536-545): per
  ! column, per band, per mode, n_layer + 1 entries bottom-up
  function prop1d(x) result(y)
    real(RealExt), intent(in) :: x(:, :, :, :)
    real(RealExt), allocatable :: y(:)
    integer :: ll
    allocate(y(n_band * n_aer * (n_layer + 1) * n_full))
    y = junk
    do l = 1, n_profile
      do ib = 1, n_band
        do jm = 1, n_aer
          do i = 1, n_layer
            ll = n_band * n_aer * (n_layer + 1) * (profile_list(l) - 1) &
               + n_aer * (n_layer + 1) * (ib - 1) + (n_layer + 1) * (jm - 1) + (n_layer + 1 - i)
            y(ll) = x(l, i, jm, ib)
          end do
        end do
      end do
    end do
  end function prop1d

  subroutine put_levels(f)
    real(RealExt), intent(in) :: f(0:, :)
    do i = 0, n_layer
      ii = n_layer - i
      write(iu) (f(ii, profile_list(l)), l = 1, n_profile)
    end do
  end subroutine put_levels

end program runes_lfric_case
