! gramsci_host.F90 -- command-line host of the MI355X N-point correlation
! engine.  It keeps the reference program's flags and output-file layout
! (src/gramsci.F90: parseOptions :1166-1273, read_files_2 :1364-1396,
! read_files :1342-1362, writers :382-397, :451-453, :513-543, :649-655,
! :829-858) and hands the
! counting to libgramsci_hip.so through the C ABI in include/gramsci_hip.h.
!
!   gramsci -gal D [-ran R] -rmin a -rmax b -nbins n [-nmu m] [-log] [-out f]
!           [-2pcf] [-3pcf] [-equi] [-4pcf] [-tetra]  |  -cut file ...  |  -help
!
! Extra option of this build: -fp64 counts with the plain FP64 kernels (the
! default is FP32 with an FP64 guard band; both give identical counts).

module gramsci_hip_api
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: GS_FP64 = 0, GS_FP32_GUARD = 1

  type, bind(C) :: gs_params
     real(c_double) :: rmin, rmax
     integer(c_int32_t) :: nbins, nmu, log_bins, precision, cell_div, reserved
  end type gs_params

  type, bind(C) :: gs_stats
     real(c_double) :: ms_grid, ms_count, ms_kernel_avg
     integer(c_int64_t) :: n_units, n_binned, n_candidates, n_recheck, n_overflow, n_evaluated
     integer(c_int32_t) :: cell_div, reserved
  end type gs_stats

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

     integer(c_int) function gs_set_catalogue(ctx, xyz, w, buf, n_data, n_rand) bind(C, name='gs_set_catalogue')
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx, xyz, w, buf
       integer(c_int64_t), value :: n_data, n_rand
     end function gs_set_catalogue

     integer(c_int) function gs_count_2pcf(ctx, p, cnt, wsum, n2, n3, st) bind(C, name='gs_count_2pcf')
       import :: c_int, c_ptr, gs_params, gs_stats
       type(c_ptr), value :: ctx
       type(gs_params), intent(in) :: p
       type(c_ptr), value :: cnt, wsum, n2, n3
       type(gs_stats), intent(out) :: st
     end function gs_count_2pcf

     integer(c_int) function gs_count_3pcf(ctx, p, cnt, wsum, n2, n3, st) bind(C, name='gs_count_3pcf')
       import :: c_int, c_ptr, gs_params, gs_stats
       type(c_ptr), value :: ctx
       type(gs_params), intent(in) :: p
       type(c_ptr), value :: cnt, wsum, n2, n3
       type(gs_stats), intent(out) :: st
     end function gs_count_3pcf

     integer(c_int) function gs_count_4pcf(ctx, p, cnt, n2, n3, disc_n, disc_r, st) bind(C, name='gs_count_4pcf')
       import :: c_int, c_ptr, gs_params, gs_stats
       type(c_ptr), value :: ctx
       type(gs_params), intent(in) :: p
       type(c_ptr), value :: cnt, n2, n3, disc_n, disc_r
       type(gs_stats), intent(out) :: st
     end function gs_count_4pcf

     integer(c_int) function gs_count_3pcf_equi(ctx, p, cnt, wsum, n2, n3, st) bind(C, name='gs_count_3pcf_equi')
       import :: c_int, c_ptr, gs_params, gs_stats
       type(c_ptr), value :: ctx
       type(gs_params), intent(in) :: p
       type(c_ptr), value :: cnt, wsum, n2, n3
       type(gs_stats), intent(out) :: st
     end function gs_count_3pcf_equi

     integer(c_int) function gs_count_4pcf_tetra(ctx, p, cnt, n2, n3, disc_n, disc_r, st) &
          bind(C, name='gs_count_4pcf_tetra')
       import :: c_int, c_ptr, gs_params, gs_stats
       type(c_ptr), value :: ctx
       type(gs_params), intent(in) :: p
       type(c_ptr), value :: cnt, n2, n3, disc_n, disc_r
       type(gs_stats), intent(out) :: st
     end function gs_count_4pcf_tetra

     type(c_ptr) function gs_last_error() bind(C, name='gs_last_error')
       import :: c_ptr
     end function gs_last_error
  end interface

contains

  function last_error() result(msg)
    character(len=1024) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ' '
    call c_f_pointer(gs_last_error(), s, [1024])
    do i = 1, 1024
       if (s(i) == c_null_char) exit
       msg(i:i) = s(i)
    end do
  end function last_error

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) then
       print '(a,": ",a)', what, trim(last_error())
       error stop 1
    end if
  end subroutine check

end module gramsci_hip_api


program gramsci
  use iso_c_binding
  use gramsci_hip_api
  implicit none

  character(len=2000) :: file_gal, file_ran, file_out
  real(c_double) :: rmin, rmax
  integer :: nbins, nmu
  logical :: use_log, have_ran, cut_mode, mpi_mode, fp64
  logical :: do2, do3, do4, do_equi, do_tetra
  integer(c_int64_t) :: n_data, n_rand
  real(c_double), allocatable, target :: xyz(:,:), w(:)
  integer(c_int32_t), allocatable, target :: buf(:)
  real(c_double), allocatable :: edges(:)
  real(c_double) :: mu_scale
  type(c_ptr) :: ctx
  type(gs_params) :: prm
  integer :: j

  call defaults()
  call parse_options()

  if (cut_mode) then
     call read_cut()
  else
     call read_pair()
  end if

  ! bin edges as printed (gramsci.F90:88-96)
  allocate(edges(nbins + 1))
  do j = 0, nbins
     if (use_log) then
        edges(j + 1) = 10**(log10(rmin) + j*(log10(rmax) - log10(rmin))/float(nbins))
     else
        edges(j + 1) = rmin + j*(rmax - rmin)/float(nbins)
     end if
  end do
  mu_scale = 0.5*nmu

  prm%rmin = rmin
  prm%rmax = rmax
  prm%nbins = nbins
  prm%nmu = nmu
  prm%log_bins = merge(1, 0, use_log)
  prm%precision = merge(GS_FP64, GS_FP32_GUARD, fp64)
  prm%cell_div = 0
  prm%reserved = 0

  call check(gs_create(ctx, 0_c_int), 'gs_create')
  if (cut_mode) then
     call check(gs_set_catalogue(ctx, c_loc(xyz), c_loc(w), c_loc(buf), n_data, n_rand), 'gs_set_catalogue')
  else
     call check(gs_set_catalogue(ctx, c_loc(xyz), c_loc(w), c_null_ptr, n_data, n_rand), 'gs_set_catalogue')
  end if

  ! each requested analysis writes the same output file, in the reference's order (gramsci.F90:193-214)
  if (do2) call run_2pcf()
  if (do3) call run_3pcf()
  if (do_equi) call run_equi()
  if (do4) call run_4pcf()
  if (do_tetra) call run_tetra()

  call check(gs_destroy(ctx), 'gs_destroy')
  print '(a)', ' Exit... stage left!'

contains

  subroutine defaults()
    file_gal = ' '
    file_ran = ' '
    file_out = 'result.txt'
    rmin = 0.0d0
    rmax = 0.0d0
    nbins = 0
    nmu = 1
    use_log = .false.
    have_ran = .false.
    cut_mode = .false.
    mpi_mode = .false.
    fp64 = .false.
    do2 = .false.; do3 = .false.; do4 = .false.; do_equi = .false.; do_tetra = .false.
  end subroutine defaults

  subroutine help()
    print '(a)', ' gramsci: binned 2-, 3- and 4-point correlation counts on an MI355X'
    print '(a)', ' '
    print '(a)', '   gramsci -gal galaxy_file [-ran random_file] -rmin Rmin -rmax Rmax -nbins Nbins'
    print '(a)', '           [-nmu Nmu] [-log] [-out out_file] [-fp64]'
    print '(a)', '           [-2pcf] [-3pcf] [-equi] [-4pcf] [-tetra]'
    print '(a)', '   gramsci -cut subsample_file ...    (columns x y z w buffer; weights not normalised)'
    print '(a)', ' '
    print '(a)', '   Input files hold x y z w per line.  Nmu >= 2 bins mu linearly in -1 < mu < 1.'
  end subroutine help

  ! gramsci.F90:1166-1273: options are matched on their first 6 characters
  subroutine parse_options()
    integer :: n, i
    character(len=6) :: opt
    character(len=2000) :: arg
    n = command_argument_count()
    if (n < 6) then
       print '(a)', ' Not enough input parameters. Please read the following help info'
       call help()
       stop
    end if
    i = 1
    do while (i <= n)
       call get_command_argument(i, opt)
       select case (opt)
       case ('-gal')
          call get_command_argument(i + 1, arg); file_gal = arg; i = i + 2
       case ('-ran')
          have_ran = .true.
          call get_command_argument(i + 1, arg); file_ran = arg; i = i + 2
       case ('-out')
          call get_command_argument(i + 1, arg); file_out = arg; i = i + 2
       case ('-rmin')
          call get_command_argument(i + 1, arg); read (arg, *) rmin; i = i + 2
       case ('-rmax')
          call get_command_argument(i + 1, arg); read (arg, *) rmax; i = i + 2
       case ('-nbins')
          call get_command_argument(i + 1, arg); read (arg, *) nbins; i = i + 2
       case ('-nmu')
          call get_command_argument(i + 1, arg); read (arg, *) nmu; i = i + 2
          if (nmu >= 2) print '(a)', ' Anisotropic analysis requested'
       case ('-cut')
          call get_command_argument(i + 1, arg)
          cut_mode = .true.
          file_gal = arg
          file_out = arg          ! the reference writes its result over the sub-sample file
          i = i + 2
       case ('-mpi')
          mpi_mode = .true.; i = i + 1
       case ('-log')
          use_log = .true.; i = i + 1
          print '(a)', ' Using logarithmically spaced bins'
       case ('-help')
          call help(); stop
       case ('-tetra')
          do_tetra = .true.; i = i + 1
       case ('-3pcf')
          do3 = .true.; i = i + 1
       case ('-4pcf')
          do4 = .true.; i = i + 1
       case ('-equi')
          do_equi = .true.; i = i + 1
       case ('-2pcf')
          do2 = .true.; i = i + 1
       case ('-fp64')
          fp64 = .true.; i = i + 1
       case default
          print '("unknown option ",a," ignored")', trim(opt)
          stop
       end select
    end do
  end subroutine parse_options

  ! number of records of a list-directed file; unreadable records are skipped
  ! (gramsci.F90:1319-1340)
  integer(c_int64_t) function count_records(fname) result(n)
    character(len=*), intent(in) :: fname
    real(c_double) :: aux
    integer :: ios
    n = 0
    open(11, file=fname, status='old', action='read')
    do
       read(11, *, iostat=ios) aux
       if (ios < 0) exit
       if (ios > 0) cycle
       n = n + 1
    end do
    close(11)
  end function count_records

  ! -gal/-ran: x y z w per record, raw weights (the engine normalises them as
  ! gramsci.F90:1390-1391 does)
  subroutine read_pair()
    integer(c_int64_t) :: i
    n_data = count_records(trim(file_gal))
    n_rand = 0
    if (have_ran) n_rand = count_records(trim(file_ran))
    print '(a,i0,a)', ' Preparing to read ', n_data, ' data points'
    print '(a,i0,a)', ' Preparing to read ', n_rand, ' random points'
    allocate(xyz(3, n_data + n_rand), w(n_data + n_rand), buf(1))
    open(11, file=trim(file_gal), status='old', action='read')
    do i = 1, n_data
       read(11, *) xyz(1:3, i), w(i)
    end do
    close(11)
    if (have_ran) then
       open(11, file=trim(file_ran), status='old', action='read')
       do i = n_data + 1, n_data + n_rand
          read(11, *) xyz(1:3, i), w(i)
       end do
       close(11)
    end if
  end subroutine read_pair

  ! -cut: x y z w buffer per record; every point is "data" and the weights are
  ! used as read (gramsci.F90:1342-1362)
  subroutine read_cut()
    integer(c_int64_t) :: i
    if (mpi_mode) file_gal = '1.loadnodes'
    n_data = count_records(trim(file_gal))
    n_rand = 0
    print '(a,i0,a)', ' Preparing to read ', n_data, ' data points'
    allocate(xyz(3, n_data), w(n_data), buf(n_data))
    open(11, file=trim(file_gal), status='old', action='read')
    do i = 1, n_data
       read(11, *) xyz(1:3, i), w(i), buf(i)
    end do
    close(11)
  end subroutine read_cut

  ! raw (un-normalised) sums of the 4PCF in -cut mode, whose entry points return the normalised
  ! sums only: N2 = sum(w_i w_j w_k w_l) over the data = normalised N2 * (sum w)^4
  real(c_double) function cut_scale(k) result(s)
    integer, intent(in) :: k
    real(c_double) :: sw
    sw = sum(w(1:n_data))
    s = sw**k
  end function cut_scale

  subroutine report(st, what)
    type(gs_stats), intent(in) :: st
    character(len=*), intent(in) :: what
    print '(1x,a," on the GPU: grid ",f10.3," ms, count ",f12.3," ms, ",i0," binned")', &
         what, st%ms_grid, st%ms_count, st%n_binned
  end subroutine report

  subroutine run_2pcf()
    integer(c_int64_t), allocatable, target :: cnt(:,:,:)
    real(c_double), allocatable, target :: n2(:,:), n3(:,:), ws(:,:,:)
    type(gs_stats) :: st
    integer :: l, k
    allocate(cnt(nmu, nbins, 3), n2(nmu, nbins), n3(nmu, nbins), ws(nmu, nbins, 3))
    call check(gs_count_2pcf(ctx, prm, c_loc(cnt), c_loc(ws), c_loc(n2), c_loc(n3), st), 'gs_count_2pcf')
    call report(st, '2PCF')
    if (cut_mode) then
       ! every point is data with its weight as read: N2 is the raw data-data sum, which the
       ! engine returns exactly (class 1 of the weight sums)
       n2 = ws(:, :, 1)
       n3 = 0.0d0
    end if
    ! gramsci.F90:382-397
    open(11, file=trim(file_out), status='unknown')
    write(11, *) 'r1 min, r1 max, r2 min, r2 max, NN, RR, 2pcf (xi)'
    do l = 1, nbins
       do k = 1, nmu
          write(11, '(8(e14.7,1x))') edges(l), edges(l + 1), ((float(k) - 1.)/mu_scale) - 1., &
               (float(k)/mu_scale) - 1.0, n2(k, l), n3(k, l), n2(k, l)/n3(k, l)
       end do
    end do
    close(11)
  end subroutine run_2pcf

  subroutine run_3pcf()
    integer(c_int64_t), allocatable, target :: cnt(:,:,:)
    real(c_double), allocatable, target :: n2(:,:), n3(:,:), ws(:,:,:)
    type(gs_stats) :: st
    integer :: i1, i2, i3, l, cfg, ncfg
    ncfg = nbins*(nbins + 1)*(nbins + 2)/6
    allocate(cnt(4, nmu, ncfg), n2(nmu, ncfg), n3(nmu, ncfg), ws(4, nmu, ncfg))
    call check(gs_count_3pcf(ctx, prm, c_loc(cnt), c_loc(ws), c_loc(n2), c_loc(n3), st), 'gs_count_3pcf')
    call report(st, '3PCF')
    if (cut_mode) then
       n2 = ws(1, :, :)      ! raw data-data-data sum (all points are data)
       n3 = 0.0d0
    end if
    ! gramsci.F90:513-543, configurations in i <= j <= k order; with nmu >= 2 one row per mu bin
    open(11, file=trim(file_out), status='unknown')
    if (nmu >= 2) then
       write(11, *) 'r1 min, r1 max, r2 min, r2 max, r3 min, r3 max, mu min, mu max, NNN, RRR, 3pcf (zeta)'
    else
       write(11, *) 'r1 min, r1 max, r2 min, r2 max, r3 min, r3 max, NNN, RRR, 3pcf (zeta)'
    end if
    cfg = 0
    do i1 = 1, nbins
       do i2 = i1, nbins
          do i3 = i2, nbins
             cfg = cfg + 1
             if (nmu >= 2) then
                do l = 1, nmu
                   write(11, '(11(e14.7,1x))') edges(i1), edges(i1 + 1), edges(i2), edges(i2 + 1), edges(i3), &
                        edges(i3 + 1), ((float(l) - 1.)/mu_scale/2.), (float(l)/mu_scale/2.), n2(l, cfg), &
                        n3(l, cfg), n2(l, cfg)/n3(l, cfg)
                end do
             else
                write(11, '(9(e14.7,1x))') edges(i1), edges(i1 + 1), edges(i2), edges(i2 + 1), edges(i3), &
                     edges(i3 + 1), n2(1, cfg), n3(1, cfg), n2(1, cfg)/n3(1, cfg)
             end if
          end do
       end do
    end do
    close(11)
  end subroutine run_3pcf

  ! -equi: triangles whose three sides share one bin (gramsci.F90:402-461); no header line
  subroutine run_equi()
    integer(c_int64_t), allocatable, target :: cnt(:,:,:)
    real(c_double), allocatable, target :: n2(:,:), n3(:,:), ws(:,:,:)
    type(gs_stats) :: st
    integer :: l, k
    allocate(cnt(4, nmu, nbins), n2(nmu, nbins), n3(nmu, nbins), ws(4, nmu, nbins))
    call check(gs_count_3pcf_equi(ctx, prm, c_loc(cnt), c_loc(ws), c_loc(n2), c_loc(n3), st), &
         'gs_count_3pcf_equi')
    call report(st, 'equilateral 3PCF')
    if (cut_mode) then
       n2 = ws(1, :, :)      ! raw data-data-data sum (all points are data)
       n3 = 0.0d0
    end if
    open(11, file=trim(file_out), status='unknown')
    do l = 1, nbins
       do k = 1, nmu
          write(11, '(8(e14.7,1x))') edges(l), edges(l + 1), ((float(k) - 1.)/mu_scale/2.), &
               (float(k)/mu_scale/2.), n2(k, l), n3(k, l), n2(k, l)/n3(k, l)
       end do
    end do
    close(11)
  end subroutine run_equi

  ! -tetra: regular tetrahedra (gramsci.F90:548-660); no header line.  The
  ! counts have no mu dependence: rows k >= 2 print zero counts, as the
  ! reference's N2(l,k,3) = 0 for k >= 2.
  subroutine run_tetra()
    integer(c_int64_t), allocatable, target :: cnt(:,:)
    real(c_double), allocatable, target :: n2(:), n3(:), dn(:,:), dr(:,:)
    real(c_double) :: a2, a3
    type(gs_params) :: p1
    type(gs_stats) :: st
    integer :: l, k
    allocate(cnt(16, nbins), n2(nbins), n3(nbins), dn(nbins, 6), dr(nbins, 6))
    p1 = prm
    p1%nmu = 1
    call check(gs_count_4pcf_tetra(ctx, p1, c_loc(cnt), c_loc(n2), c_loc(n3), c_loc(dn), c_loc(dr), st), &
         'gs_count_4pcf_tetra')
    call report(st, 'tetrahedral 4PCF')
    if (cut_mode) then
       n2 = n2*cut_scale(4)
       n3 = 0.0d0
       dn = dn*cut_scale(3)
       dr = 0.0d0
    end if
    open(11, file=trim(file_out), status='unknown')
    do l = 1, nbins
       do k = 1, nmu
          a2 = 0.0d0
          a3 = 0.0d0
          if (k == 1) then
             a2 = n2(l)
             a3 = n3(l)
          end if
          write(11, '(9(e14.7,1x))') edges(l), edges(l + 1), ((float(k) - 1.)/mu_scale/2.), &
               (float(k)/mu_scale/2.), a2, a3, a2/a3, &
               (dn(l, 1)/dr(l, 1))*(dn(l, 5)/dr(l, 5)) + (dn(l, 3)/dr(l, 3))*(dn(l, 4)/dr(l, 4)) &
               + (dn(l, 2)/dr(l, 2))*(dn(l, 6)/dr(l, 6))
       end do
    end do
    close(11)
  end subroutine run_tetra

  subroutine run_4pcf()
    integer(c_int64_t), allocatable, target :: cnt(:,:,:,:)
    real(c_double), allocatable, target :: n2(:,:,:), n3(:,:,:), dn(:,:,:,:), dr(:,:,:,:)
    type(gs_stats) :: st
    integer :: i1, i2, i3, i4, k1, k2, cfg, ncfg
    ncfg = nbins*(nbins + 1)*(nbins + 2)*(nbins + 3)/24
    allocate(cnt(16, nbins, nbins, ncfg), n2(nbins, nbins, ncfg), n3(nbins, nbins, ncfg))
    allocate(dn(nbins, nbins, ncfg, 6), dr(nbins, nbins, ncfg, 6))
    call check(gs_count_4pcf(ctx, prm, c_loc(cnt), c_loc(n2), c_loc(n3), c_loc(dn), c_loc(dr), st), &
         'gs_count_4pcf')
    call report(st, '4PCF')
    if (cut_mode) then
       n2 = n2*cut_scale(4)
       n3 = 0.0d0
       dn = dn*cut_scale(2)
       dr = 0.0d0
    end if
    ! gramsci.F90:829-858: perimeter configurations in i <= j <= k <= l order, then (ind2, ind6)
    open(11, file=trim(file_out), status='unknown')
    write(11, *) 'r1 min, r1 max, r2 min, r2 max, r3 min, r3 max, r4 min, r4 max, NNNN, RRRR, 4pcf (eta)'
    cfg = 0
    do i1 = 1, nbins
       do i2 = i1, nbins
          do i3 = i2, nbins
             do i4 = i3, nbins
                cfg = cfg + 1
                do k1 = 1, nbins
                   do k2 = 1, nbins
                      write(11, '(17(e14.7,1x))') edges(i1), edges(i1 + 1), edges(i2), edges(i2 + 1), &
                           edges(i3), edges(i3 + 1), edges(i4), edges(i4 + 1), &
                           n2(k2, k1, cfg), n3(k2, k1, cfg), (n2(k2, k1, cfg)/n3(k2, k1, cfg)), &
                           (dn(k2, k1, cfg, 1)/dr(k2, k1, cfg, 1))*(dn(k2, k1, cfg, 5)/dr(k2, k1, cfg, 5)), &
                           (dn(k2, k1, cfg, 3)/dr(k2, k1, cfg, 3))*(dn(k2, k1, cfg, 4)/dr(k2, k1, cfg, 4)), &
                           (dn(k2, k1, cfg, 2)/dr(k2, k1, cfg, 2))*(dn(k2, k1, cfg, 6)/dr(k2, k1, cfg, 6))
                   end do
                end do
             end do
          end do
       end do
    end do
    close(11)
  end subroutine run_4pcf

end program gramsci
