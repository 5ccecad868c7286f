! Test-infrastructure harness: calls the reference McICA generator
! (socrates_cloud_gen::cloud_gen) on inputs read from a binary file and
! writes the sub-column condensate scalings c_sub and the number of cloudy
! sub-columns. Used only to write golden vectors for the parity tests.
!
! Input (stream, native): int32 n_profile, n_layer, n_subcol, ipph, ioverlap,
! n1, n2; int32 rand_seed(n_profile); float64 rlc_cf, rlc_cw, sigma_qcw,
! avg_cf, zf as (n_profile, n_layer); float64 xcw(n1, n2).
! Output: int32 n_subcol_cld(n_profile); float64 c_sub(n_profile, n_layer, n_subcol).
!
! Usage: soc_cloud_gen_case <in.bin> <out.bin>
program soc_cloud_gen_case
  use socrates_cloud_gen, only: cloud_gen
  use realtype_rd, only: RealK
  implicit none
  character(len=1024) :: fin, fout
  integer :: iu, np, nl, ns, ipph, iov, n1, n2
  integer, allocatable :: seed(:), ncld(:)
  real(RealK), allocatable :: rlc_cf(:,:), rlc_cw(:,:), sig(:,:), cf(:,:), zf(:,:), &
    xcw(:,:), c_sub(:,:,:), c_cloud(:,:), c_ratio(:,:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=iu, file=trim(fin), access='stream', form='unformatted', status='old')
  read(iu) np, nl, ns, ipph, iov, n1, n2
  allocate(seed(np), ncld(np), rlc_cf(np, nl), rlc_cw(np, nl), sig(np, nl), cf(np, nl), &
           zf(np, nl), xcw(n1, n2), c_sub(np, nl, ns), c_cloud(np, nl), c_ratio(np, nl))
  read(iu) seed
  read(iu) rlc_cf, rlc_cw, sig, cf, zf
  read(iu) xcw
  close(iu)
  c_cloud = 0.0_RealK
  c_ratio = 0.0_RealK
  c_sub = 0.0_RealK
  call cloud_gen(nl, 1, nl, np, 1, np, ns, n1, n2, ipph, iov, seed, &
                 rlc_cf, rlc_cw, sig, cf, c_cloud, c_ratio, zf, xcw, ncld, c_sub)
  open(newunit=iu, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(iu) ncld
  write(iu) c_sub
  close(iu)
end program soc_cloud_gen_case
