! domain_decomposition.F90 -- drop-in for the reference's domain decomposition
! tool (src/domain_decomposition.f90).  Usage, as the tool:
!
!   domain_decomposition RMAX        (reads ./allpoints, writes ./<r>.loadnodes)
!
! "allpoints" holds x y z w gal aux region per record; every record that does
! not read is skipped (:15-33).  For every region label r = 1..max(region)
! the file "<r>.loadnodes" receives, in input order, each point strictly
! inside the bounding box of region r grown by RMAX, as "x y z w gal buffer"
! with buffer = 0 inside the closed box and 1 outside (:38-74).  Coordinates,
! weights and RMAX are default reals, as in the tool; the boxes, memberships
! and buffer flags are computed on the GPU by gs_decompose (libgramsci_hip.so).
!
! The tool sizes its arrays one record larger than the file and writes that
! never-read last element like a point; this program writes read points only.

module gramsci_decompose_api
  use iso_c_binding
  implicit none
  interface
     integer(c_int) function gs_create(ctx, device) bind(C, name='gs_create')
       import :: c_int, c_ptr
       type(c_ptr), intent(out) :: ctx
       integer(c_int), value :: device
     end function gs_create

     integer(c_int) function gs_destroy(ctx) bind(C, name='gs_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function gs_destroy

     integer(c_int) function gs_decompose(ctx, xyz, region, n, rmax, nreg, off, idx, buf, cap, box) &
          bind(C, name='gs_decompose')
       import :: c_int, c_ptr, c_int64_t, c_float
       type(c_ptr), value :: ctx, xyz, region, off, idx, buf, box
       integer(c_int64_t), value :: n, cap
       real(c_float), value :: rmax
       integer(c_int), value :: nreg
     end function gs_decompose

     type(c_ptr) function gs_last_error() bind(C, name='gs_last_error')
       import :: c_ptr
     end function gs_last_error
  end interface

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: s(:)
    character(len=1024) :: msg
    integer :: i
    if (rc == 0) return
    msg = ' '
    call c_f_pointer(gs_last_error(), s, [1024])
    do i = 1, 1024
       if (s(i) == c_null_char) exit
       msg(i:i) = s(i)
    end do
    print '(a,": ",a)', what, trim(msg)
    error stop 1
  end subroutine check

end module gramsci_decompose_api


program domain_decomposition
  use iso_c_binding
  use gramsci_decompose_api
  implicit none

  real(c_float) :: rmax, aux
  real(c_float), allocatable, target :: xyz(:,:), box(:,:)
  real, allocatable :: wgt(:)
  integer(c_int32_t), allocatable, target :: gal(:), region(:), idx(:), buf(:)
  integer(c_int64_t), allocatable, target :: off(:)
  integer(c_int64_t) :: n, i, k
  integer :: ios, nreg, r
  character(len=100) :: c
  type(c_ptr) :: ctx

  call get_command_argument(1, c)
  read (c, *) rmax
  print *, rmax

  ! records readable as one real (domain_decomposition.f90:15-20)
  n = 0
  open (11, file='allpoints', status='unknown')
  do
     read (11, *, iostat=ios) aux
     if (ios < 0) exit
     if (ios > 0) cycle
     n = n + 1
  end do
  close (11)

  allocate (xyz(3, max(n, 1_c_int64_t)), wgt(max(n, 1_c_int64_t)), gal(max(n, 1_c_int64_t)), &
            region(max(n, 1_c_int64_t)))
  ! x y z w gal aux region per point; a record that fails to read is skipped (:25-33)
  i = 0
  open (11, file='allpoints', status='unknown')
  do while (i < n)
     read (11, *, iostat=ios) xyz(1:3, i + 1), wgt(i + 1), gal(i + 1), aux, region(i + 1)
     if (ios < 0) exit
     if (ios > 0) cycle
     i = i + 1
  end do
  close (11)
  n = i

  nreg = 0
  if (n > 0) nreg = maxval(region(1:n))
  print *, "creating ", nreg, " regions"
  if (nreg < 1) stop

  call check(gs_create(ctx, 0_c_int), 'gs_create')
  allocate (off(nreg + 1), box(6, nreg))
  call check(gs_decompose(ctx, c_loc(xyz), c_loc(region), n, rmax, int(nreg, c_int), c_loc(off), c_null_ptr, &
                          c_null_ptr, 0_c_int64_t, c_loc(box)), 'gs_decompose')
  allocate (idx(max(off(nreg + 1), 1_c_int64_t)), buf(max(off(nreg + 1), 1_c_int64_t)))
  call check(gs_decompose(ctx, c_loc(xyz), c_loc(region), n, rmax, int(nreg, c_int), c_loc(off), c_loc(idx), &
                          c_loc(buf), off(nreg + 1), c_null_ptr), 'gs_decompose')
  call check(gs_destroy(ctx), 'gs_destroy')

  do r = 1, nreg
     print *, box(1, r), box(4, r), box(2, r), box(5, r), box(3, r), box(6, r)
     print *, box(1, r) - rmax, box(4, r) + rmax, box(2, r) - rmax, box(5, r) + rmax, box(3, r) - rmax, &
          box(6, r) + rmax
     write (c, "(I3,A10)") r, ".loadnodes"
     c = adjustl(c)
     print *, c
     open (unit=11, file=trim(c), status='unknown')
     do k = off(r) + 1, off(r + 1)
        i = idx(k) + 1
        write (11, *) xyz(:, i), wgt(i), gal(i), buf(k)
     end do
     close (11)
  end do

end program domain_decomposition
