! example_chain4.f90 -- a reference-style Fortran driver using the drop-in.
!
! The 4-site Hubbard chain of drivers/vca_hm_chain1d.f90 (U=4, half filling
! with HFMODE=T and XMU=0, t = t' = -1, Nk k-points on the reduced zone),
! written against module vca_gpu instead of module VCA: the calls are the
! reference's vca_init_solver / vca_solve / vca_get_sft_potential and the
! grid getters vca_get_sigma_matsubara / vca_get_gimp_realaxis, and the
! pole file of save_gfprime (gfprime_f.restart, unformatted).
program example_chain4
  use vca_gpu
  implicit none
  integer, parameter :: L = 4, nk = 16
  complex(8), parameter :: xi = (0d0, 1d0)
  complex(8) :: hloc(L, L, 1, 1, 1, 1), hk(L, L, 1, 1, 1, 1, nk)
  complex(8), allocatable :: smats(:,:,:,:,:,:,:), greal(:,:,:,:,:)
  real(8) :: sft, kpt, pi
  integer :: i, ik

  pi = acos(-1d0)
  Nlat = L
  Uloc(1) = 4d0
  beta = 100d0
  lanc_method = "lanczos"
  hloc = (0d0, 0d0)
  hk = (0d0, 0d0)
  do i = 1, L - 1
     hloc(i, i + 1, 1, 1, 1, 1) = -1d0
     hloc(i + 1, i, 1, 1, 1, 1) = -1d0
  end do
  do ik = 1, nk
     kpt = 2d0 * pi * (ik - 1) / (nk * L)
     hk(:, :, :, :, :, :, ik) = hloc
     hk(1, L, 1, 1, 1, 1, ik) = hk(1, L, 1, 1, 1, 1, ik) - exp(xi * kpt * L)
     hk(L, 1, 1, 1, 1, 1, ik) = hk(L, 1, 1, 1, 1, 1, ik) - exp(-xi * kpt * L)
  end do

  call vca_init_solver()
  call vca_solve(hloc, hk)
  call vca_get_sft_potential(sft)
  write(*, "(A,ES25.17)") "SFT_POTENTIAL ", sft
  allocate(smats(Nlat, Nlat, Nspin, Nspin, Norb, Norb, Lmats))
  allocate(greal(Nspin, Nspin, Norb, Norb, Lreal))
  call vca_get_sigma_matsubara(smats)
  call vca_get_gimp_realaxis(greal, 1, 2)
  write(*, "(A,2ES25.17)") "SIGMA_MATS_11_1 ", smats(1, 1, 1, 1, 1, 1, 1)
  write(*, "(A,2ES25.17)") "SIGMA_MATS_23_7 ", smats(2, 3, 1, 1, 1, 1, 7)
  write(*, "(A,2ES25.17)") "GREAL_12_2500 ", greal(1, 1, 1, 1, 2500)
  call save_gfprime("gfprime_f")
  call vca_finalize_solver()
end program example_chain4
