!  vca_gpu.f90 -- Fortran drop-in for the VCA solver core on MI355X.
!
!  Presents the reference's public names (QcmPlab/VCA, module VCA):
!    vca_init_solver([comm,] [bath_h, bath_v])    VCA_MAIN.f90:51-115
!    vca_solve(Hloc, Hk [, bath_h, bath_v])       VCA_MAIN.f90:124-215
!    vca_get_sft_potential(sft)                   VCA_IO/get_sft_potential.f90:1-4
!  the operator plugin point, for reference-style solvers that drive H x v
!  themselves:
!    cc_sparse_HxV (abstract interface) and the procedure pointer
!    spHtimesV_p                                  VCA_VARS_GLOBAL.f90:83-89,137
!    build_Hv_sector(isector) / delete_Hv_sector() /
!    vecDim_Hv_sector(isector)                    VCA_HAMILTONIAN.f90:38-217
!    vca_set_Hcluster(Hloc), vca_set_bath(h, v)   VCA_AUX_FUNX.f90:124-131,
!                                                 VCA_BATH_SETUP.f90:112-138
!  and the input variables of VCA_INPUT_VARS.f90:14-60 under the same
!  names, and calls libvca_gpu.so through its C ABI (include/vca_gpu.h)
!  with iso_c_binding.  A nonzero status from the library stops the
!  program with the library's message, as the reference stops on errors.
!  The MPI form of vca_init_solver/vca_solve accepts a communicator and
!  ignores it: one GPU holds the whole sector.
module vca_gpu
  use iso_c_binding
  implicit none
  private

  ! ---- input variables (reference names and defaults) -----------------
  integer, public :: Nlat = 1, Norb = 1, Nspin = 1, Nlat_bath = 0, Norb_bath = 0
  real(8), public :: Uloc(5) = [2d0, 0d0, 0d0, 0d0, 0d0]
  real(8), public :: Ust = 0d0, Jh = 0d0, xmu = 0d0, beta = 1000d0
  logical, public :: hfmode = .true., hfshift = .true., vca_gf_symmetric = .false.
  character(len=16), public :: lanc_method = "arpack"
  integer, public :: lanc_nstates_sector = 6, lanc_niter = 512, lanc_ngfiter = 200
  integer, public :: lanc_dim_threshold = 256
  real(8), public :: lanc_tolerance = 1d-12, gs_threshold = 1d-9
  integer, public :: lanc_nstates_total = 1, lanc_nstates_step = 2, lanc_niter_spectrum = 10
  integer, public :: lanc_ncv_factor = 3, lanc_ncv_add = 5
  real(8), public :: cutoff = 1d-9, bandwidth = 0d0
  integer, public :: Lmats = 5000, Lreal = 5000
  real(8), public :: wini = -5d0, wfin = 5d0, eps = 0.01d0
  ! additions of this build
  integer, public :: vca_precision = 0        ! 0 FP64, 1 MIXED
  integer, public :: vca_omega_mode = 0       ! 0 adaptive quadrature, 1 Matsubara grid
  integer, public :: vca_omega_nfreq = 4096
  real(8), public :: vca_quad_epsabs = 1d-10, vca_quad_epsrel = 1d-12
  integer, allocatable, public :: vca_sectors_list(:,:)   ! (2, n): (nup, ndw)

  type, bind(C) :: vca_params_t
     integer(c_int32_t) :: nlat, norb, nspin, nlat_bath, norb_bath
     integer(c_int32_t) :: hfmode, hfshift, gf_symmetric
     real(c_double)     :: uloc(5), ust, jh, xmu, beta
     integer(c_int32_t) :: lanc_method, lanc_nstates_sector, lanc_niter, lanc_ngfiter
     integer(c_int32_t) :: lanc_dim_threshold
     real(c_double)     :: lanc_tolerance, gs_threshold
     integer(c_int32_t) :: precision, omega_mode, omega_nfreq
     real(c_double)     :: quad_epsabs, quad_epsrel
     integer(c_int32_t) :: nsectors
     type(c_ptr)        :: sectors
     integer(c_int32_t) :: lanc_nstates_total, lanc_nstates_step, lanc_niter_spectrum
     integer(c_int32_t) :: lanc_ncv_factor, lanc_ncv_add
     real(c_double)     :: cutoff, bandwidth
     integer(c_int32_t) :: lmats, lreal
     real(c_double)     :: wini, wfin, eps
  end type vca_params_t

  ! vca_results of include/vca_gpu.h (same field order and kinds)
  type, bind(C) :: vca_results_t
     real(c_double)     :: sft_potential, omegaprime, omega_integral, e0_mixed
     integer(c_int32_t) :: nstates, nchains
     integer(c_int64_t) :: hxv_count
     real(c_double)     :: hxv_states, hxv_seconds, t_diag, t_gf, t_omega
     integer(c_int64_t) :: omega_evals
     real(c_double)     :: omega_seconds, zeta, max_exc
     integer(c_int32_t) :: finite_t, nmax, spectrum_iters, nchains_derived
     integer(c_int64_t) :: npoles, grid_pole_evals
     real(c_double)     :: grid_seconds
  end type vca_results_t

  interface
     subroutine c_default_params(p) bind(C, name="vca_default_params")
       import :: vca_params_t
       type(vca_params_t), intent(inout) :: p
     end subroutine c_default_params
     function c_last_error() bind(C, name="vca_last_error") result(msg)
       import :: c_ptr
       type(c_ptr) :: msg
     end function c_last_error
     function c_init_solver(p, out) bind(C, name="vca_init_solver") result(rc)
       import :: vca_params_t, c_ptr, c_int
       type(vca_params_t), intent(in) :: p
       type(c_ptr), intent(out) :: out
       integer(c_int) :: rc
     end function c_init_solver
     subroutine c_finalize_solver(s) bind(C, name="vca_finalize_solver")
       import :: c_ptr
       type(c_ptr), value :: s
     end subroutine c_finalize_solver
     function c_solve(s, hloc, hk, nk, bath_h, bath_v) bind(C, name="vca_solve") result(rc)
       import :: c_ptr, c_int, c_int64_t
       type(c_ptr), value :: s, hloc, hk, bath_h, bath_v
       integer(c_int64_t), value :: nk
       integer(c_int) :: rc
     end function c_solve
     function c_get_sft(s, out) bind(C, name="vca_get_sft_potential") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: s
       real(c_double), intent(out) :: out
       integer(c_int) :: rc
     end function c_get_sft
     function c_get_results(s, out) bind(C, name="vca_get_results") result(rc)
       import :: c_ptr, c_int, vca_results_t
       type(c_ptr), value :: s
       type(vca_results_t), intent(out) :: out
       integer(c_int) :: rc
     end function c_get_results
     function c_save_gfprime(s, file, used, fmt) bind(C, name="vca_save_gfprime") result(rc)
       import :: c_ptr, c_int, c_int32_t, c_char
       type(c_ptr), value :: s
       character(kind=c_char), dimension(*), intent(in) :: file
       integer(c_int32_t), value :: used, fmt
       integer(c_int) :: rc
     end function c_save_gfprime
     function c_read_gfprime(s, file, used, fmt) bind(C, name="vca_read_gfprime") result(rc)
       import :: c_ptr, c_int, c_int32_t, c_char
       type(c_ptr), value :: s
       character(kind=c_char), dimension(*), intent(in) :: file
       integer(c_int32_t), value :: used, fmt
       integer(c_int) :: rc
     end function c_read_gfprime
     function c_get_gf_grid(s, axis, g, sg, g0) bind(C, name="vca_get_gf_grid") result(rc)
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: s, g, sg, g0
       integer(c_int32_t), value :: axis
       integer(c_int) :: rc
     end function c_get_gf_grid
  end interface

  ! The operator plugin point (VCA_VARS_GLOBAL.f90:83-89,137): solvers call
  ! spHtimesV_p(Nloc, v, Hv) on the sector build_Hv_sector made current.
  abstract interface
     subroutine cc_sparse_HxV(Nloc, v, Hv)
       integer :: Nloc
       complex(8), dimension(Nloc) :: v
       complex(8), dimension(Nloc) :: Hv
     end subroutine cc_sparse_HxV
  end interface
  procedure(cc_sparse_HxV), pointer, public :: spHtimesV_p => null()

  interface
     function c_set_hcluster(s, hloc) bind(C, name="vca_set_hcluster") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: s, hloc
       integer(c_int) :: rc
     end function c_set_hcluster
     function c_set_bath(s, bh, bv) bind(C, name="vca_set_bath") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: s, bh, bv
       integer(c_int) :: rc
     end function c_set_bath
     function c_sector_create(s, nup, ndw, prec, out) bind(C, name="vca_sector_create") &
          result(rc)
       import :: c_ptr, c_int, c_int32_t
       type(c_ptr), value :: s
       integer(c_int32_t), value :: nup, ndw, prec
       type(c_ptr), intent(out) :: out
       integer(c_int) :: rc
     end function c_sector_create
     subroutine c_sector_destroy(sec) bind(C, name="vca_sector_destroy")
       import :: c_ptr
       type(c_ptr), value :: sec
     end subroutine c_sector_destroy
     function c_sector_dims(sec, dim, dimup, dimdw) bind(C, name="vca_sector_dims") result(rc)
       import :: c_ptr, c_int, c_int32_t, c_int64_t
       type(c_ptr), value :: sec
       integer(c_int64_t), intent(out) :: dim
       integer(c_int32_t), intent(out) :: dimup, dimdw
       integer(c_int) :: rc
     end function c_sector_dims
     function c_hxv(sec, n, v, hv) bind(C, name="vca_hxv") result(rc)
       import :: c_ptr, c_int, c_int64_t, c_double_complex
       type(c_ptr), value :: sec
       integer(c_int64_t), value :: n
       complex(c_double_complex), intent(in) :: v(*)
       complex(c_double_complex), intent(out) :: hv(*)
       integer(c_int) :: rc
     end function c_hxv
  end interface

  type(c_ptr), save :: handle = c_null_ptr
  type(c_ptr), save :: sector_handle = c_null_ptr
  integer, save :: Hsector = 0
  type(vca_params_t), save, target :: cparams
  integer(c_int32_t), allocatable, target, save :: csectors(:)
  real(8), save :: last_omegaprime = 0d0, last_sft = 0d0

  interface vca_init_solver
     module procedure vca_init_solver_serial
     module procedure vca_init_solver_comm
  end interface vca_init_solver

  interface vca_solve
     module procedure vca_solve_serial
     module procedure vca_solve_comm
  end interface vca_solve

  ! VCA_IO.f90:15-36
  interface vca_get_gimp_matsubara
     module procedure vca_get_gimp_matsubara_full
     module procedure vca_get_gimp_matsubara_ij
  end interface vca_get_gimp_matsubara
  interface vca_get_gimp_realaxis
     module procedure vca_get_gimp_real_full
     module procedure vca_get_gimp_real_ij
  end interface vca_get_gimp_realaxis
  interface vca_get_sigma_matsubara
     module procedure vca_get_sigma_matsubara_full
     module procedure vca_get_sigma_matsubara_ij
  end interface vca_get_sigma_matsubara
  interface vca_get_sigma_realaxis
     module procedure vca_get_sigma_real_full
     module procedure vca_get_sigma_real_ij
  end interface vca_get_sigma_realaxis

  public :: vca_init_solver, vca_solve, vca_get_sft_potential, vca_finalize_solver
  public :: save_gfprime, read_gfprime
  public :: vca_get_gimp_matsubara, vca_get_gimp_realaxis
  public :: vca_get_sigma_matsubara, vca_get_sigma_realaxis
  public :: vca_params_t, vca_results_t
  public :: cc_sparse_HxV, build_Hv_sector, delete_Hv_sector, vecDim_Hv_sector, vca_HxV
  public :: vca_set_Hcluster, vca_set_bath

contains

  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == 0) return
    call c_f_pointer(c_last_error(), msg, [512])
    text = ""
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       text(i:i) = msg(i)
    end do
    write(*, "(A,A,A,I0,A,A)") "STOP: ", where, " failed (", rc, "): ", trim(text)
    stop 1
  end subroutine check

  subroutine fill_params()
    integer :: i
    call c_default_params(cparams)
    cparams%nlat = Nlat; cparams%norb = Norb; cparams%nspin = Nspin
    cparams%nlat_bath = Nlat_bath; cparams%norb_bath = Norb_bath
    cparams%hfmode = merge(1, 0, hfmode); cparams%hfshift = merge(1, 0, hfshift)
    cparams%gf_symmetric = merge(1, 0, vca_gf_symmetric)
    cparams%uloc = Uloc; cparams%ust = Ust; cparams%jh = Jh
    cparams%xmu = xmu; cparams%beta = beta
    cparams%lanc_method = merge(0, 1, trim(lanc_method) == "lanczos")
    cparams%lanc_nstates_sector = lanc_nstates_sector
    cparams%lanc_niter = lanc_niter; cparams%lanc_ngfiter = lanc_ngfiter
    cparams%lanc_dim_threshold = lanc_dim_threshold
    cparams%lanc_tolerance = lanc_tolerance; cparams%gs_threshold = gs_threshold
    cparams%precision = vca_precision; cparams%omega_mode = vca_omega_mode
    cparams%omega_nfreq = vca_omega_nfreq
    cparams%quad_epsabs = vca_quad_epsabs; cparams%quad_epsrel = vca_quad_epsrel
    cparams%lanc_nstates_total = lanc_nstates_total
    cparams%lanc_nstates_step = lanc_nstates_step
    cparams%lanc_niter_spectrum = lanc_niter_spectrum
    cparams%lanc_ncv_factor = lanc_ncv_factor; cparams%lanc_ncv_add = lanc_ncv_add
    cparams%cutoff = cutoff; cparams%bandwidth = bandwidth
    cparams%lmats = Lmats; cparams%lreal = Lreal
    cparams%wini = wini; cparams%wfin = wfin; cparams%eps = eps
    cparams%nsectors = 0; cparams%sectors = c_null_ptr
    if (allocated(vca_sectors_list)) then
       if (allocated(csectors)) deallocate(csectors)
       allocate(csectors(2 * size(vca_sectors_list, 2)))
       do i = 1, size(vca_sectors_list, 2)
          csectors(2 * i - 1) = vca_sectors_list(1, i)
          csectors(2 * i) = vca_sectors_list(2, i)
       end do
       cparams%nsectors = size(vca_sectors_list, 2)
       cparams%sectors = c_loc(csectors)
    end if
  end subroutine fill_params

  ! VCA_MAIN.f90:51-76.  With bath_h and bath_v the bath dimensions must be
  ! positive (the reference stops otherwise, :55-56); without them the
  ! cluster has no bath: Nlat_bath = Norb_bath = 0 (:59-62).  The arrays
  ! themselves are the model input of vca_solve / vca_set_bath, as in the
  ! reference, where init only allocates the bath structure.
  subroutine vca_init_solver_serial(bath_h, bath_v)
    complex(8), intent(in), optional :: bath_h(:,:,:,:,:,:), bath_v(:,:,:,:,:,:)
    if (present(bath_h) .neqv. present(bath_v)) stop "vca_init_solver: bath_h and bath_v go together"
    if (present(bath_h) .and. present(bath_v)) then
       if (Nlat_bath < 1 .or. Norb_bath < 1) stop "VCA_INIT_SOLVER error: wrong bath dimensions"
    else
       write(*, "(A)") "Bath not present, setting Nbath to 0"
       Nlat_bath = 0
       Norb_bath = 0
    end if
    if (c_associated(handle)) return   ! dimensions are fixed once, as isetup does
    call fill_params()
    call check(c_init_solver(cparams, handle), "vca_init_solver")
  end subroutine vca_init_solver_serial

  subroutine vca_init_solver_comm(comm, bath_h, bath_v)
    integer, intent(in) :: comm
    complex(8), intent(in), optional :: bath_h(:,:,:,:,:,:), bath_v(:,:,:,:,:,:)
    if (comm < -huge(comm)) stop
    call vca_init_solver_serial(bath_h, bath_v)
  end subroutine vca_init_solver_comm

  ! VCA_MAIN.f90:124-215
  subroutine vca_solve_serial(Hloc, Hk, bath_h, bath_v)
    complex(8), intent(in), target, contiguous :: Hloc(:,:,:,:,:,:)
    complex(8), intent(in), target, contiguous :: Hk(:,:,:,:,:,:,:)
    complex(8), intent(inout), target, contiguous, optional :: bath_h(:,:,:,:,:,:)
    complex(8), intent(inout), target, contiguous, optional :: bath_v(:,:,:,:,:,:)
    type(c_ptr) :: pbh, pbv
    integer :: unit
    real(8) :: sft
    type(vca_results_t) :: res
    if (.not. c_associated(handle)) call vca_init_solver_serial()
    if (any(shape(Hloc) /= [Nlat, Nlat, Nspin, Nspin, Norb, Norb])) &
         stop "STOP: wrong cluster matrix dimensions"
    if (any(shape(Hk(:,:,:,:,:,:,1)) /= [Nlat, Nlat, Nspin, Nspin, Norb, Norb])) &
         stop "STOP: wrong lattice matrix dimensions"
    pbh = c_null_ptr
    pbv = c_null_ptr
    if (present(bath_h) .and. present(bath_v)) then
       if (any(shape(bath_h) /= [Nlat_bath, Nlat_bath, Nspin, Nspin, Norb_bath, Norb_bath])) &
            stop "vca_solve: wrong bath_h shape"
       if (any(shape(bath_v) /= [Nlat, Nlat_bath, Nspin, Nspin, Norb, Norb_bath])) &
            stop "vca_solve: wrong bath_v shape"
       pbh = c_loc(bath_h)
       pbv = c_loc(bath_v)
    end if
    call check(c_solve(handle, c_loc(Hloc), c_loc(Hk), int(size(Hk, 7), c_int64_t), pbh, pbv), &
         "vca_solve")
    call check(c_get_sft(handle, sft), "vca_get_sft_potential")
    call check(c_get_results(handle, res), "vca_get_results")
    last_sft = sft
    last_omegaprime = res%omegaprime
    ! SFT_potential.vca, VCA_MAIN.f90:199-203 (sft, omegaprime, -integral)
    open(newunit=unit, file="SFT_potential.vca", position="append")
    write(unit, *) sft, res%omegaprime, -res%omega_integral
    close(unit)
  end subroutine vca_solve_serial

  subroutine vca_solve_comm(comm, Hloc, Hk, bath_h, bath_v)
    integer, intent(in) :: comm
    complex(8), intent(in), target, contiguous :: Hloc(:,:,:,:,:,:)
    complex(8), intent(in), target, contiguous :: Hk(:,:,:,:,:,:,:)
    complex(8), intent(inout), target, contiguous, optional :: bath_h(:,:,:,:,:,:)
    complex(8), intent(inout), target, contiguous, optional :: bath_v(:,:,:,:,:,:)
    if (comm < -huge(comm)) stop
    call vca_solve_serial(Hloc, Hk, bath_h, bath_v)
  end subroutine vca_solve_comm

  ! VCA_IO/get_sft_potential.f90:1-4
  subroutine vca_get_sft_potential(sft)
    real(8), intent(out) :: sft
    call check(c_get_sft(handle, sft), "vca_get_sft_potential")
  end subroutine vca_get_sft_potential

  ! VCA_AUX_FUNX.f90:608-652 and 654-697: file//".restart" (".used")
  subroutine save_gfprime(file, used, use_formatted)
    character(len=*), intent(in), optional :: file
    logical, intent(in), optional :: used, use_formatted
    call check(c_save_gfprime(handle, gf_file(file), gf_flag(used), gf_flag(use_formatted)), &
         "save_gfprime")
  end subroutine save_gfprime

  subroutine read_gfprime(file, used, use_formatted)
    character(len=*), intent(in), optional :: file
    logical, intent(in), optional :: used, use_formatted
    if (.not. c_associated(handle)) call vca_init_solver_serial()
    call check(c_read_gfprime(handle, gf_file(file), gf_flag(used), gf_flag(use_formatted)), &
         "read_gfprime")
  end subroutine read_gfprime

  function gf_file(file) result(c)
    character(len=*), intent(in), optional :: file
    character(kind=c_char, len=:), allocatable :: c
    if (present(file)) then
       c = trim(file) // c_null_char
    else
       c = "gfprime" // c_null_char
    end if
  end function gf_file

  integer(c_int32_t) function gf_flag(x)
    logical, intent(in), optional :: x
    gf_flag = 0
    if (present(x)) then
       if (x) gf_flag = 1
    end if
  end function gf_flag

  ! One array (which: 1 G, 2 Sigma, 3 G0) on the Matsubara (axis 0) or real
  ! (axis 1) grid of the solver, [Nlat,Nlat,Nspin,Nspin,Norb,Norb,L].
  subroutine grid_full(axis, which, A, name)
    integer, intent(in) :: axis, which
    complex(8), intent(inout), target, contiguous :: A(:,:,:,:,:,:,:)
    character(len=*), intent(in) :: name
    type(c_ptr) :: p(3)
    integer :: L
    if (.not. c_associated(handle)) stop "vca_get: solver not initialised"
    L = cparams%lmats
    if (axis == 1) L = cparams%lreal
    if (any(shape(A) /= [Nlat, Nlat, Nspin, Nspin, Norb, Norb, L])) then
       write(*, "(A)") name // ": wrong array shape"
       stop 1
    end if
    p = c_null_ptr
    p(which) = c_loc(A)
    call check(c_get_gf_grid(handle, int(axis, c_int32_t), p(1), p(2), p(3)), name)
  end subroutine grid_full

  subroutine grid_ij(axis, which, A, ilat, jlat, name)
    integer, intent(in) :: axis, which, ilat, jlat
    complex(8), intent(inout) :: A(:,:,:,:,:)
    character(len=*), intent(in) :: name
    complex(8), allocatable, target :: full(:,:,:,:,:,:,:)
    integer :: L
    L = cparams%lmats
    if (axis == 1) L = cparams%lreal
    allocate(full(Nlat, Nlat, Nspin, Nspin, Norb, Norb, L))
    call grid_full(axis, which, full, name)
    A = full(ilat, jlat, :, :, :, :, :)
  end subroutine grid_ij

  ! VCA_IO/get_gimp_matsubara.f90
  subroutine vca_get_gimp_matsubara_full(Gmats)
    complex(8), intent(inout), target, contiguous :: Gmats(:,:,:,:,:,:,:)
    call grid_full(0, 1, Gmats, "vca_get_gimp_matsubara")
  end subroutine vca_get_gimp_matsubara_full
  subroutine vca_get_gimp_matsubara_ij(Gmats, ilat, jlat)
    complex(8), intent(inout) :: Gmats(:,:,:,:,:)
    integer, intent(in) :: ilat, jlat
    call grid_ij(0, 1, Gmats, ilat, jlat, "vca_get_gimp_matsubara")
  end subroutine vca_get_gimp_matsubara_ij
  ! VCA_IO/get_gimp_realaxis.f90
  subroutine vca_get_gimp_real_full(Greal)
    complex(8), intent(inout), target, contiguous :: Greal(:,:,:,:,:,:,:)
    call grid_full(1, 1, Greal, "vca_get_gimp_realaxis")
  end subroutine vca_get_gimp_real_full
  subroutine vca_get_gimp_real_ij(Greal, ilat, jlat)
    complex(8), intent(inout) :: Greal(:,:,:,:,:)
    integer, intent(in) :: ilat, jlat
    call grid_ij(1, 1, Greal, ilat, jlat, "vca_get_gimp_realaxis")
  end subroutine vca_get_gimp_real_ij
  ! VCA_IO/get_sigma_matsubara.f90
  subroutine vca_get_sigma_matsubara_full(Smats)
    complex(8), intent(inout), target, contiguous :: Smats(:,:,:,:,:,:,:)
    call grid_full(0, 2, Smats, "vca_get_sigma_matsubara")
  end subroutine vca_get_sigma_matsubara_full
  subroutine vca_get_sigma_matsubara_ij(Smats, ilat, jlat)
    complex(8), intent(inout) :: Smats(:,:,:,:,:)
    integer, intent(in) :: ilat, jlat
    call grid_ij(0, 2, Smats, ilat, jlat, "vca_get_sigma_matsubara")
  end subroutine vca_get_sigma_matsubara_ij
  ! VCA_IO/get_sigma_realaxis.f90
  subroutine vca_get_sigma_real_full(Sreal)
    complex(8), intent(inout), target, contiguous :: Sreal(:,:,:,:,:,:,:)
    call grid_full(1, 2, Sreal, "vca_get_sigma_realaxis")
  end subroutine vca_get_sigma_real_full
  subroutine vca_get_sigma_real_ij(Sreal, ilat, jlat)
    complex(8), intent(inout) :: Sreal(:,:,:,:,:)
    integer, intent(in) :: ilat, jlat
    call grid_ij(1, 2, Sreal, ilat, jlat, "vca_get_sigma_realaxis")
  end subroutine vca_get_sigma_real_ij

  ! set_Hcluster_nnn, VCA_AUX_FUNX.f90:124-131
  subroutine vca_set_Hcluster(Hloc)
    complex(8), intent(in), target, contiguous :: Hloc(:,:,:,:,:,:)
    if (.not. c_associated(handle)) stop "vca_set_Hcluster: solver not initialised"
    if (any(shape(Hloc) /= [Nlat, Nlat, Nspin, Nspin, Norb, Norb])) &
         stop "set_Hcluster_nnn: wrong Hloc shape"
    call check(c_set_hcluster(handle, c_loc(Hloc)), "vca_set_Hcluster")
  end subroutine vca_set_Hcluster

  ! VCA_BATH_SETUP.f90:112-138 (the shape checks of :117-118)
  subroutine vca_set_bath(h_in, v_in)
    complex(8), intent(in), target, contiguous :: h_in(:,:,:,:,:,:), v_in(:,:,:,:,:,:)
    if (.not. c_associated(handle)) stop "vca_set_bath: solver not initialised"
    if (any(shape(h_in) /= [Nlat_bath, Nlat_bath, Nspin, Nspin, Norb_bath, Norb_bath])) &
         stop "vca_set_bath: wrong bath_h shape"
    if (any(shape(v_in) /= [Nlat, Nlat_bath, Nspin, Nspin, Norb, Norb_bath])) &
         stop "vca_set_bath: wrong bath_v shape"
    call check(c_set_bath(handle, c_loc(h_in), c_loc(v_in)), "vca_set_bath")
  end subroutine vca_set_bath

  ! Sector numbering of get_Sector / get_Nup / get_Ndw (VCA_SETUP.f90:404-450)
  ! with one up and one down string: isector = 1 + Ndw + Nup (Ns + 1).
  subroutine sector_n(isector, nup, ndw)
    integer, intent(in) :: isector
    integer, intent(out) :: nup, ndw
    integer :: ns
    ns = Nlat * Norb + Nlat_bath * Norb_bath
    if (isector < 1 .or. isector > (ns + 1)**2) stop "build_Hv_sector: isector out of range"
    ndw = mod(isector - 1, ns + 1)
    nup = (isector - 1) / (ns + 1)
  end subroutine sector_n

  ! build_Hv_sector, VCA_HAMILTONIAN.f90:38-146: makes sector isector of the
  ! current Hloc (and bath) the operator's sector and points spHtimesV_p at
  ! its H x v; the hop tables are built on the GPU.  The precision is
  ! vca_precision.
  subroutine build_Hv_sector(isector)
    integer, intent(in) :: isector
    integer :: nup, ndw
    if (.not. c_associated(handle)) stop "build_Hv_sector: solver not initialised"
    call sector_n(isector, nup, ndw)
    if (c_associated(sector_handle)) call delete_Hv_sector()
    call check(c_sector_create(handle, int(nup, c_int32_t), int(ndw, c_int32_t), &
         int(vca_precision, c_int32_t), sector_handle), "build_Hv_sector")
    Hsector = isector
    spHtimesV_p => vca_HxV
  end subroutine build_Hv_sector

  ! delete_Hv_sector, VCA_HAMILTONIAN.f90:149-186
  subroutine delete_Hv_sector()
    if (c_associated(sector_handle)) call c_sector_destroy(sector_handle)
    sector_handle = c_null_ptr
    Hsector = 0
    spHtimesV_p => null()
  end subroutine delete_Hv_sector

  ! vecDim_Hv_sector, VCA_HAMILTONIAN.f90:193-217 (one GPU: the whole sector)
  integer function vecDim_Hv_sector(isector) result(vecDim)
    integer, intent(in) :: isector
    integer :: nup, ndw, ns
    call sector_n(isector, nup, ndw)
    ns = Nlat * Norb + Nlat_bath * Norb_bath
    vecDim = binom(ns, nup) * binom(ns, ndw)
  end function vecDim_Hv_sector

  integer function binom(n, k)
    integer, intent(in) :: n, k
    integer :: i
    binom = 1
    do i = 1, k
       binom = binom * (n - k + i) / i
    end do
  end function binom

  ! The operator of the current sector with the cc_sparse_HxV interface
  ! (VCA_VARS_GLOBAL.f90:83-89): Hv = H v, state i = iup + (idw-1) DimUp
  ! (1-based), the product on the GPU (vca_hxv).  It is what
  ! build_Hv_sector binds spHtimesV_p to, in place of directMatVec_main
  ! (VCA_HAMILTONIAN_DIRECT_HxV.f90:42-78).
  subroutine vca_HxV(Nloc, v, Hv)
    integer :: Nloc
    complex(8), dimension(Nloc) :: v
    complex(8), dimension(Nloc) :: Hv
    integer(c_int64_t) :: dim
    integer(c_int32_t) :: du, dd
    if (.not. c_associated(sector_handle)) stop "vca_HxV: no sector (build_Hv_sector first)"
    call check(c_sector_dims(sector_handle, dim, du, dd), "vca_HxV")
    if (int(Nloc, c_int64_t) /= dim) stop "vca_HxV: Nloc differs from the sector dimension"
    call check(c_hxv(sector_handle, dim, v, Hv), "vca_HxV")
  end subroutine vca_HxV

  subroutine vca_finalize_solver()
    if (c_associated(sector_handle)) call delete_Hv_sector()
    if (c_associated(handle)) call c_finalize_solver(handle)
    handle = c_null_ptr
  end subroutine vca_finalize_solver

end module vca_gpu
