! gfprime_writer.f90 -- writes a cluster Green's function pole file the
! way the reference's save_gfprime does (VCA_AUX_FUNX.f90:484-604, 608-652):
! a derived type with a defined WRITE(UNFORMATTED) whose child transfers put
! Nstates, per state Nchan, per channel Nexc, poles, weights into the record
! the parent WRITE opens, one parent WRITE per element in the order ilat,
! jlat, ispin, iorb, jorb.  Built and run by make_gfprime_golden.py with the
! image's Fortran compiler, so the record framing comes from a Fortran
! runtime rather than from this project's writers.
!
! Contents (Nlat=2, Nspin=1, Norb=1; all values exact in binary):
!   element k = 1..4, Nstates = 2, Nchan = 2 (diagonal) or 4 (off-diagonal),
!   Nexc = mod(k + s + c, 4) (so some channels are empty),
!   pole(j)   = k + s/4 + c/16 + j/256            (imaginary part 0)
!   weight(j) = (j/8, -(s + c)/32)
module golden_gf
  implicit none
  type spectrum
     complex(8), allocatable :: weight(:), poles(:)
  end type spectrum
  type chanlist
     type(spectrum), allocatable :: channel(:)
  end type chanlist
  type gfm
     type(chanlist), allocatable :: state(:)
  end type gfm
  interface write(unformatted)
     procedure gfm_write
  end interface write(unformatted)
contains
  subroutine gfm_write(dtv, unit, iostat, iomsg)
    class(gfm), intent(in) :: dtv
    integer, intent(in) :: unit
    integer, intent(out) :: iostat
    character(*), intent(inout) :: iomsg
    integer :: s, c
    write(unit, iostat=iostat, iomsg=iomsg) size(dtv%state)
    do s = 1, size(dtv%state)
       write(unit, iostat=iostat, iomsg=iomsg) size(dtv%state(s)%channel)
       do c = 1, size(dtv%state(s)%channel)
          write(unit, iostat=iostat, iomsg=iomsg) size(dtv%state(s)%channel(c)%poles), &
               dtv%state(s)%channel(c)%poles, dtv%state(s)%channel(c)%weight
       end do
    end do
  end subroutine gfm_write
end module golden_gf

program gfprime_writer
  use golden_gf
  implicit none
  type(gfm) :: g(4)
  integer :: k, s, c, j, n, il, jl
  k = 0
  do il = 1, 2
     do jl = 1, 2
        k = k + 1
        allocate(g(k)%state(2))
        do s = 1, 2
           if (il == jl) then
              allocate(g(k)%state(s)%channel(2))
           else
              allocate(g(k)%state(s)%channel(4))
           end if
           do c = 1, size(g(k)%state(s)%channel)
              n = mod(k + s + c, 4)
              allocate(g(k)%state(s)%channel(c)%poles(n), g(k)%state(s)%channel(c)%weight(n))
              do j = 1, n
                 g(k)%state(s)%channel(c)%poles(j) = &
                      cmplx(dble(k) + dble(s) / 4d0 + dble(c) / 16d0 + dble(j) / 256d0, 0d0, 8)
                 g(k)%state(s)%channel(c)%weight(j) = &
                      cmplx(dble(j) / 8d0, -dble(s + c) / 32d0, 8)
              end do
           end do
        end do
     end do
  end do
  open(10, file="gfprime_flang.restart", form="unformatted", access="sequential", &
       status="replace")
  do k = 1, 4
     write(10) g(k)
  end do
  close(10)
end program gfprime_writer
