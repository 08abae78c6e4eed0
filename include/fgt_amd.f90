! fgt_amd.f90 -- type-checked Fortran interface to libfgt_amd (iso_c_binding).
!
! fgt_boxfgt takes the arguments of the reference boxfgt
! (src/bfgt/boxfgt.f:5-8) and returns a status (0 = done); callers that keep
! the reference's implicit interface call boxfgt (symbol boxfgt_) instead,
! which the library exports with the same arguments and no status.
module fgt_amd
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function fgt_boxfgt(nd,ndim,delta,eps,ipoly,iperiod,     &
        norder,npbox,nboxes,nlevels,ltree,itree,iptr,centers,boxsize,      &
        fvals,ifpgh,pot,grad,hess,ifnewtree,ntarg,targs,ifpghtarg,pote,    &
        grade,hesse,timeinfo) bind(C, name="fgt_boxfgt")
      import :: c_int, c_double
      integer(c_int) :: nd, ndim, ipoly, iperiod, norder, npbox, nboxes
      integer(c_int) :: nlevels, ltree, ifpgh, ifnewtree, ntarg, ifpghtarg
      real(c_double) :: delta, eps
      integer(c_int) :: itree(*), iptr(8)
      real(c_double) :: centers(*), boxsize(*), fvals(*), pot(*), grad(*)
      real(c_double) :: hess(*), targs(*), pote(*), grade(*), hesse(*)
      real(c_double) :: timeinfo(*)
    end function
    integer(c_int) function fgt_last_status(msg, len) bind(C, name="fgt_last_status")
      import :: c_int, c_char
      character(kind=c_char) :: msg(*)
      integer(c_int), value :: len
    end function
  end interface
end module
