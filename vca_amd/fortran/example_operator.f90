! example_operator.f90 -- a reference-style solver loop on the operator
! plugin point of the drop-in.
!
! A caller that drives H x v itself, as SciFortran's sp_lanc_tridiag does
! through spHtimesV_p (VCA_GF_NORMAL.f90:218-223; interface cc_sparse_HxV,
! VCA_VARS_GLOBAL.f90:83-89): it reads a cluster with a bath and a start
! vector from operator_in.bin (stream: Hloc, bath_h, bath_v as complex(8)
! arrays of the module's shapes, then v0 of the sector), makes the sector
! current with build_Hv_sector, runs 20 steps of the plain Lanczos
! recurrence through spHtimesV_p, prints alpha and beta, and writes H v0 to
! operator_hv.bin.
program example_operator
  use vca_gpu
  implicit none
  integer, parameter :: nsteps = 20
  complex(8), allocatable :: hloc(:,:,:,:,:,:), bath_h(:,:,:,:,:,:), bath_v(:,:,:,:,:,:)
  complex(8), allocatable :: v(:), vold(:), w(:)
  real(8) :: alpha(nsteps), betas(nsteps), nrm
  integer :: u, nloc, isector, nup, ndw, k, ns

  Nlat = 3
  Nlat_bath = 3
  Norb_bath = 1
  Uloc(1) = 3d0
  hfshift = .false.
  ns = Nlat + Nlat_bath
  nup = 3
  ndw = 2
  allocate(hloc(Nlat, Nlat, 1, 1, 1, 1), bath_h(Nlat_bath, Nlat_bath, 1, 1, 1, 1))
  allocate(bath_v(Nlat, Nlat_bath, 1, 1, 1, 1))
  isector = 1 + ndw + nup * (ns + 1)
  call vca_init_solver(bath_h, bath_v)
  nloc = vecDim_Hv_sector(isector)
  allocate(v(nloc), vold(nloc), w(nloc))
  open(newunit=u, file="operator_in.bin", access="stream", form="unformatted", status="old")
  read(u) hloc, bath_h, bath_v, v
  close(u)
  call vca_set_bath(bath_h, bath_v)
  call vca_set_Hcluster(hloc)
  call build_Hv_sector(isector)
  call spHtimesV_p(nloc, v, w)
  open(newunit=u, file="operator_hv.bin", access="stream", form="unformatted", status="replace")
  write(u) w
  close(u)
  ! plain Lanczos from the normalised v0 (the sp_lanc_tridiag recurrence)
  v = v / sqrt(real(dot_product(v, v), 8))
  vold = (0d0, 0d0)
  betas(1) = 0d0
  do k = 1, nsteps
     call spHtimesV_p(nloc, v, w)
     alpha(k) = real(dot_product(v, w), 8)
     w = w - alpha(k) * v
     if (k > 1) w = w - betas(k) * vold
     if (k == nsteps) exit
     nrm = sqrt(real(dot_product(w, w), 8))
     betas(k + 1) = nrm
     vold = v
     v = w / nrm
  end do
  do k = 1, nsteps
     write(*, "(A,I0,2ES25.17)") "AB ", k, alpha(k), betas(k)
  end do
  call delete_Hv_sector()
  call vca_finalize_solver()
end program example_operator
