! A Fortran caller of the drop-in, written the way the reference's own test
! driver calls boxfgt (test/bfgt/test_boxfgt_all.f:366-369): implicit
! interface, every argument by reference, pot zeroed by the caller.
!
! usage: drop_in_driver <input.bin> <output.bin> [ifnewtree]
! input (stream, little-endian): int32 nd, ndim, ipoly, iperiod, norder,
! npbox, nboxes, nlevels, ltree; float64 delta, eps; int32 itree(ltree),
! iptr(8); float64 centers(ndim*nboxes), boxsize(nlevels+1),
! fvals(nd*npbox*nboxes).  output: float64 pot(nd*npbox*nboxes).
! With ifnewtree /= 0 the driver calls fgt_boxfgt through the fgt_amd module
! and prints the status and message instead (tests pass an invalid value).
program drop_in_driver
  use fgt_amd
  implicit none
  integer :: nd, ndim, ipoly, iperiod, norder, npbox, nboxes, nlevels, ltree
  integer :: ifpgh, ifnewtree, ntarg, ifpghtarg, st, i
  double precision :: delta, eps
  integer, allocatable :: itree(:)
  integer :: iptr(8)
  double precision, allocatable :: centers(:), boxsize(:), fvals(:), pot(:)
  double precision :: grad(1), hess(1), targs(1), pote(1), grade(1), hesse(1), timeinfo(20)
  character(len=512) :: fin, fout, arg
  character(kind=c_char) :: msg(256)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  ifnewtree = 0
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    read(arg, *) ifnewtree
  end if

  open(unit=10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) nd, ndim, ipoly, iperiod, norder, npbox, nboxes, nlevels, ltree
  read(10) delta, eps
  allocate(itree(ltree), centers(ndim*nboxes), boxsize(nlevels+1))
  allocate(fvals(nd*npbox*nboxes), pot(nd*npbox*nboxes))
  read(10) itree, iptr, centers, boxsize, fvals
  close(10)

  pot = 0
  ifpgh = 1
  ntarg = 0
  ifpghtarg = 0
  if (ifnewtree == 0) then
    call boxfgt(nd, ndim, delta, eps, ipoly, iperiod, norder, npbox,      &
        nboxes, nlevels, ltree, itree, iptr, centers, boxsize, fvals,     &
        ifpgh, pot, grad, hess, ifnewtree, ntarg, targs, ifpghtarg,       &
        pote, grade, hesse, timeinfo)
    st = fgt_last_status(msg, 256)
    print '(a,i0)', 'status ', st
    open(unit=11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(11) pot
    close(11)
  else
    st = fgt_boxfgt(nd, ndim, delta, eps, ipoly, iperiod, norder, npbox,  &
        nboxes, nlevels, ltree, itree, iptr, centers, boxsize, fvals,     &
        ifpgh, pot, grad, hess, ifnewtree, ntarg, targs, ifpghtarg,       &
        pote, grade, hesse, timeinfo)
    print '(a,i0)', 'status ', st
    st = fgt_last_status(msg, 256)
    do i = 1, 256
      if (msg(i) == c_null_char) exit
    end do
    print '(a,256a)', 'message ', msg(1:i-1)
    print '(a,es12.4)', 'pot max ', maxval(abs(pot))
  end if
end program
