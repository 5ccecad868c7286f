! Example host program of the Fortran binding: the same calls a host model
! makes through the reference interface (set_spectrum, set_mcica, runes with
! a caller-associated StrDiag), here on one case read from a binary file
! written by socrates_amd/case_io.py. Writes flux_up, flux_down, flux_direct
! and heating_rate in binary.
!
! Usage: runes_gpu_case <sw_spectral_file> <lw_spectral_file> <mcica_file|-> \
!                       <case.bin> <out.bin> [fp64|mixed]
program runes_gpu_case

  use socrates_set_spectrum, only: set_spectrum, set_mcica
  use socrates_runes, only: runes, StrDiag, ip_source_illuminate, &
                            ip_precision_fp64, ip_precision_mixed
  use realtype_rd, only: RealExt

  implicit none

  character(len=1024) :: sw_file, lw_file, mcica_file, case_file, out_file, arg
  character(len=2) :: spec_name
  integer :: iu, n_profile, n_layer, i_source, l_cloud, i_overlap, l_rayleigh
  integer :: magic, version, n_band, i_st_water, i_st_ice, i_cloud_repr
  integer :: i_inhom, i_mcica_sampling, n_subcol_gen, i_solver_kind, prec
  real(RealExt) :: co2_mix_ratio, cloud_vertical_decorr, cloud_horizontal_rsd
  real(RealExt) :: conv_vertical_decorr
  ! well-mixed n2o, ch4, o2, so2, ocs, cfc11, cfc12, cfc113, hcfc22, hfc134a
  real(RealExt) :: wm(10)
  logical :: inv, last
  real(RealExt), allocatable, dimension(:,:) :: p_layer, t_layer, mass, density, &
    layer_heat_capacity, h2o, o3, p_level, t_level, albedo_diff, albedo_dir, &
    cloud_frac, liq_mmr, liq_dim, conv_frac, liq_conv_mmr, liq_conv_dim
  real(RealExt), allocatable, dimension(:) :: t_ground, cos_zen, solar_irrad
  integer, allocatable :: rand_seed(:)
  real(RealExt), allocatable, target, dimension(:,:) :: flux_up, flux_down, &
    flux_direct, heating_rate
  type(StrDiag) :: diag

  call get_command_argument(1, sw_file)
  call get_command_argument(2, lw_file)
  call get_command_argument(3, mcica_file)
  call get_command_argument(4, case_file)
  call get_command_argument(5, out_file)
  prec = ip_precision_fp64
  if (command_argument_count() >= 6) then
    call get_command_argument(6, arg)
    if (trim(arg) == 'mixed') prec = ip_precision_mixed
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
    liq_conv_dim(n_profile, n_layer), &
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
  close(iu)
  if (i_solver_kind /= 0) stop 'triple-column cases are not on this binding'
  ! optional 7th argument 'invert': hand runes bottom-up profiles
  ! (l_invert = .true.) and flip its outputs back for the output file
  ! or 'last': fields as (levels or layers or bands, n_profile) with
  ! l_profile_last = .true., outputs transposed back
  inv = .false.
  last = .false.
  if (command_argument_count() >= 7) then
    call get_command_argument(7, arg)
    inv = trim(arg) == 'invert'
    last = trim(arg) == 'last'
  end if
  if (last) then
    p_layer = transpose(p_layer); t_layer = transpose(t_layer)
    mass = transpose(mass); density = transpose(density)
    layer_heat_capacity = transpose(layer_heat_capacity)
    h2o = transpose(h2o); o3 = transpose(o3)
    p_level = transpose(p_level); t_level = transpose(t_level)
    albedo_diff = transpose(albedo_diff); albedo_dir = transpose(albedo_dir)
    cloud_frac = transpose(cloud_frac); liq_mmr = transpose(liq_mmr)
    liq_dim = transpose(liq_dim)
  end if
  if (inv) then
    p_layer = p_layer(:, n_layer:1:-1); t_layer = t_layer(:, n_layer:1:-1)
    mass = mass(:, n_layer:1:-1); density = density(:, n_layer:1:-1)
    layer_heat_capacity = layer_heat_capacity(:, n_layer:1:-1)
    h2o = h2o(:, n_layer:1:-1); o3 = o3(:, n_layer:1:-1)
    p_level = p_level(:, n_layer:0:-1); t_level = t_level(:, n_layer:0:-1)
    cloud_frac = cloud_frac(:, n_layer:1:-1); liq_mmr = liq_mmr(:, n_layer:1:-1)
    liq_dim = liq_dim(:, n_layer:1:-1)
  end if

  if (last) then
    allocate(flux_up(0:n_layer, n_profile), flux_down(0:n_layer, n_profile), &
      flux_direct(0:n_layer, n_profile), heating_rate(n_layer, n_profile))
  else
    allocate(flux_up(n_profile, 0:n_layer), flux_down(n_profile, 0:n_layer), &
      flux_direct(n_profile, 0:n_layer), heating_rate(n_profile, n_layer))
  end if
  flux_direct = 0.0_RealExt
  diag%flux_up => flux_up
  diag%flux_down => flux_down
  diag%heating_rate => heating_rate
  if (i_source == ip_source_illuminate) diag%flux_direct => flux_direct

  call set_spectrum(spectrum_name='sw', spectral_file=trim(sw_file), &
                    l_all_gases=.true.)
  call set_spectrum(spectrum_name='lw', spectral_file=trim(lw_file), &
                    l_all_gases=.true.)
  if (trim(mcica_file) /= '-') then
    call set_mcica(trim(mcica_file), sw_spectrum_name='sw', &
                   lw_spectrum_name='lw')
  end if
  if (i_source == ip_source_illuminate) then
    spec_name = 'sw'
  else
    spec_name = 'lw'
  end if

  if (l_cloud == 0) then
    call runes(n_profile, n_layer, diag, spectrum_name=spec_name, &
      i_source=i_source, p_layer=p_layer, t_layer=t_layer, &
      p_level=p_level, t_level=t_level, mass=mass, density=density, &
      h2o=h2o, o3=o3, co2_mix_ratio=co2_mix_ratio, &
      n2o_mix_ratio=wm(1), ch4_mix_ratio=wm(2), o2_mix_ratio=wm(3), &
      so2_mix_ratio=wm(4), ocs_mix_ratio=wm(5), cfc11_mix_ratio=wm(6), &
      cfc12_mix_ratio=wm(7), cfc113_mix_ratio=wm(8), hcfc22_mix_ratio=wm(9), &
      hfc134a_mix_ratio=wm(10), &
      t_ground=t_ground, cos_zenith_angle=cos_zen, solar_irrad=solar_irrad, &
      albedo_diff=albedo_diff, albedo_dir=albedo_dir, &
      l_rayleigh=(l_rayleigh /= 0), layer_heat_capacity=layer_heat_capacity, &
      l_invert=inv, l_profile_last=last, precision=prec)
  else
    call runes(n_profile, n_layer, diag, spectrum_name=spec_name, &
      i_source=i_source, p_layer=p_layer, t_layer=t_layer, &
      p_level=p_level, t_level=t_level, mass=mass, density=density, &
      h2o=h2o, o3=o3, co2_mix_ratio=co2_mix_ratio, &
      n2o_mix_ratio=wm(1), ch4_mix_ratio=wm(2), o2_mix_ratio=wm(3), &
      so2_mix_ratio=wm(4), ocs_mix_ratio=wm(5), cfc11_mix_ratio=wm(6), &
      cfc12_mix_ratio=wm(7), cfc113_mix_ratio=wm(8), hcfc22_mix_ratio=wm(9), &
      hfc134a_mix_ratio=wm(10), &
      t_ground=t_ground, cos_zenith_angle=cos_zen, solar_irrad=solar_irrad, &
      albedo_diff=albedo_diff, albedo_dir=albedo_dir, &
      l_rayleigh=(l_rayleigh /= 0), layer_heat_capacity=layer_heat_capacity, &
      cloud_frac=cloud_frac, liq_mmr=liq_mmr, liq_dim=liq_dim, &
      cloud_vertical_decorr=cloud_vertical_decorr, &
      cloud_horizontal_rsd=cloud_horizontal_rsd, &
      i_cloud_representation=i_cloud_repr, i_overlap=i_overlap, &
      i_inhom=i_inhom, i_mcica_sampling=i_mcica_sampling, &
      i_st_water=i_st_water, n_subcol_gen=n_subcol_gen, rand_seed=rand_seed, &
      l_invert=inv, l_profile_last=last, precision=prec)
  end if
  write(*, '(a)') 'runes done'
  if (last) then
    flux_up = transpose(flux_up); flux_down = transpose(flux_down)
    flux_direct = transpose(flux_direct); heating_rate = transpose(heating_rate)
  end if
  if (inv) then
    flux_up = flux_up(:, n_layer:0:-1); flux_down = flux_down(:, n_layer:0:-1)
    flux_direct = flux_direct(:, n_layer:0:-1); heating_rate = heating_rate(:, n_layer:1:-1)
  end if

  open(newunit=iu, file=trim(out_file), access='stream', form='unformatted', &
       status='replace')
  write(iu) flux_up, flux_down, flux_direct, heating_rate
  close(iu)

end program runes_gpu_case
