! Test-infrastructure harness: calls the reference two_stream
! (two_stream.F90; two_coeff + ir_source + column_solver with
! solver_homogen_direct) on prescribed layer optical properties read from a
! binary file, and writes the direct flux and the interleaved total flux.
! Used only to check the oracle's kernel-level restatement.
!
! Input (stream, native): int32 n_profile, n_layer, isolir, l_scale_solar,
! l_ir_source_quad; float64 (n_profile, n_layer) tau, omega, asymmetry;
! (n_profile) sec_0, flux_inc_down, flux_inc_direct, diffuse_albedo,
! direct_albedo; (n_profile, n_layer) diff_planck, diff_planck_2;
! (n_profile) d_planck_flux_surface; (n_profile, n_layer) adjust_solar_ke.
! Output: float64 flux_direct(n_profile, 0:n_layer),
! flux_total(n_profile, 2*n_layer+2).
!
! Usage: soc_two_stream_case <in.bin> <out.bin>
program soc_two_stream_case
  use two_stream_mod, only: two_stream
  use def_control, only: StrCtrl
  use def_bound, only: StrBound
  use def_spherical_geometry, only: StrSphGeo
  use rad_pcf, only: ip_solar, ip_pifm80, ip_elsasser, ip_solver_homogen_direct, &
                     ip_scatter_full
  use realtype_rd, only: RealK
  implicit none
  character(len=1024) :: fin, fout
  integer :: iu, np, nl, isolir, lss, lq, ierr, i2s
  real(RealK), allocatable :: tau(:,:), omega(:,:), asym(:,:), sec0(:), finc(:), &
    fdir0(:), adiff(:), adir(:), dpl(:,:), dpl2(:,:), dsurf(:), adj(:,:), &
    fdir(:,:), ftot(:,:)
  type(StrCtrl) :: control
  type(StrBound) :: bound
  type(StrSphGeo) :: sph

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=iu, file=trim(fin), access='stream', form='unformatted', status='old')
  read(iu) np, nl, isolir, lss, lq
  allocate(tau(np, nl), omega(np, nl), asym(np, nl), sec0(np), finc(np), fdir0(np), &
           adiff(np), adir(np), dpl(np, nl), dpl2(np, nl), dsurf(np), adj(np, nl), &
           fdir(np, 0:nl), ftot(np, 2*nl+2))
  read(iu) tau, omega, asym
  read(iu) sec0, finc, fdir0, adiff, adir
  read(iu) dpl, dpl2
  read(iu) dsurf
  read(iu) adj
  close(iu)

  if (isolir == ip_solar) then
    i2s = ip_pifm80
  else
    i2s = ip_elsasser
  end if
  control%isolir = isolir
  control%i_2stream = i2s
  control%l_ir_source_quad = (lq /= 0)
  allocate(sph%common%path_div(np, nl))
  sph%common%path_div = 0.0_RealK
  ierr = 0
  fdir = 0.0_RealK
  ftot = 0.0_RealK
  call two_stream(ierr, control, bound, np, nl, i2s, ip_solver_homogen_direct, &
    ip_scatter_full, (lss /= 0), adj, isolir, dpl, (lq /= 0), dpl2, &
    finc, fdir0, sec0, adiff, adir, dsurf, sph, tau, tau, omega, asym, &
    fdir, ftot, np, nl, 2)
  open(newunit=iu, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(iu) fdir
  write(iu) ftot
  close(iu)
end program soc_two_stream_case
