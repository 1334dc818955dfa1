! Fortran driver of the HIP boxfgt drop-in, in the style of the reference's
! known-answer test hpots/test_initpot.f: a Fourier mode on a uniform
! periodic quadtree goes through `call boxfgt(...)` with the reference's
! 28-argument interface; pot/grad/hess are compared with the closed form of
! the periodic heat semigroup, pi*delta*exp(-delta|k|^2/4) f.
! Exits with status 1 when the relative error exceeds the tolerance.
program test_boxfgt_dropin
  implicit none
  integer, parameter :: nlev = 4, n = 16, nd = 1
  integer :: nboxes, ltree, iptr(8), nb, l, ibox, j, i, k, ip, it, ntest
  integer, allocatable :: itree(:)
  real(8), allocatable :: centers(:,:), fvals(:,:,:), pot(:,:,:), grad(:,:,:,:), hess(:,:,:,:)
  real(8) :: boxsize(0:nlev), xq(n), targs(2,1), pote(1), grade(2), hesse(3), timeinfo(20)
  real(8) :: pi, kx, ky, phi, delta, eps, s, x, y, arg, ex(6), err(3), nrm(3), deltas(3), tol
  integer :: sx, sy
  external :: boxfgt

  pi = 4*atan(1.0d0)
  nboxes = (4**(nlev+1) - 1)/3
  ltree = 2*(nlev+1) + 17*nboxes
  allocate(itree(ltree), centers(2,nboxes), fvals(nd,n*n,nboxes), pot(nd,n*n,nboxes))
  allocate(grad(nd,2,n*n,nboxes), hess(nd,3,n*n,nboxes))
  iptr(1) = 1
  iptr(2) = 2*(nlev+1) + 1
  iptr(3) = iptr(2) + nboxes
  iptr(4) = iptr(3) + nboxes
  iptr(5) = iptr(4) + nboxes
  iptr(6) = iptr(5) + 4*nboxes
  iptr(7) = iptr(6) + nboxes
  iptr(8) = iptr(7) + 9*nboxes
  itree = 0
  do l = 0, nlev
    boxsize(l) = 1.0d0/2**l
  end do
  ! root, then children of every box in parent order: mm, pm, mp, pp
  itree(1) = 1
  itree(2) = 1
  centers(:,1) = 0
  itree(iptr(2)) = 0
  itree(iptr(3)) = -1
  itree(iptr(5):iptr(5)+3) = -1
  nb = 1
  do l = 0, nlev-1
    do ibox = itree(2*l+1), itree(2*l+2)
      itree(iptr(4)+ibox-1) = 4
      do j = 1, 4
        nb = nb + 1
        sx = 1
        sy = 1
        if (j == 1 .or. j == 3) sx = -1
        if (j <= 2) sy = -1
        centers(1,nb) = centers(1,ibox) + sx*boxsize(l+1)/2
        centers(2,nb) = centers(2,ibox) + sy*boxsize(l+1)/2
        itree(iptr(5)+4*(ibox-1)+j-1) = nb
        itree(iptr(3)+nb-1) = ibox
        itree(iptr(2)+nb-1) = l+1
        itree(iptr(4)+nb-1) = 0
        itree(iptr(5)+4*(nb-1):iptr(5)+4*(nb-1)+3) = -1
      end do
    end do
    itree(2*(l+1)+1) = itree(2*l+2) + 1
    itree(2*(l+1)+2) = nb
  end do

  call gauss_nodes(n, xq)
  kx = 2*pi*2
  ky = -2*pi*1
  phi = 0.7d0
  do ibox = 1, nboxes
    do j = 1, n
      do i = 1, n
        x = centers(1,ibox) + xq(i)*boxsize(itree(iptr(2)+ibox-1))/2
        y = centers(2,ibox) + xq(j)*boxsize(itree(iptr(2)+ibox-1))/2
        fvals(1, i+n*(j-1), ibox) = cos(kx*x + ky*y + phi)
      end do
    end do
  end do

  deltas = (/ 1.0d-1, 1.0d-3, 1.0d-4 /)
  eps = 1.0d-9
  tol = 1.0d-7
  ntest = 0
  ! each delta twice: the second round reuses the plans the first one built
  do it = 1, 6
    delta = deltas(mod(it - 1, 3) + 1)
    pot = 0
    grad = 0
    hess = 0
    call boxfgt(nd, 2, delta, eps, 0, 1, n, n*n, nboxes, nlev, ltree, itree, iptr, centers, &
        boxsize, fvals, 3, pot, grad, hess, 0, 0, targs, 0, pote, grade, hesse, timeinfo)
    err = 0
    nrm = 0
    s = pi*delta*exp(-delta*(kx*kx + ky*ky)/4)
    do ibox = itree(2*nlev+1), itree(2*nlev+2)
      do j = 1, n
        do i = 1, n
          ip = i + n*(j-1)
          x = centers(1,ibox) + xq(i)*boxsize(nlev)/2
          y = centers(2,ibox) + xq(j)*boxsize(nlev)/2
          arg = kx*x + ky*y + phi
          ex(1) = s*cos(arg)
          ex(2) = -s*sin(arg)*kx
          ex(3) = -s*sin(arg)*ky
          ex(4) = -s*cos(arg)*kx*kx
          ex(5) = -s*cos(arg)*kx*ky
          ex(6) = -s*cos(arg)*ky*ky
          err(1) = err(1) + (pot(1,ip,ibox) - ex(1))**2
          nrm(1) = nrm(1) + ex(1)**2
          do k = 1, 2
            err(2) = err(2) + (grad(1,k,ip,ibox) - ex(1+k))**2
            nrm(2) = nrm(2) + ex(1+k)**2
          end do
          do k = 1, 3
            err(3) = err(3) + (hess(1,k,ip,ibox) - ex(3+k))**2
            nrm(3) = nrm(3) + ex(3+k)**2
          end do
        end do
      end do
    end do
    err = sqrt(err/nrm)
    write(*,'(a,es9.2,a,3es10.2,a,f9.6,a)') 'delta=', delta, ' rel err pot/grad/hess:', err, &
        '  (', sum(timeinfo(1:7)), ' s)'
    if (maxval(err) > tol) ntest = ntest + 1
  end do
  call boxfgt_release_plans()
  if (ntest > 0) then
    write(*,*) 'FAIL'
    error stop 1
  end if
  write(*,*) 'PASS'

contains

  subroutine gauss_nodes(m, xs)
    integer, intent(in) :: m
    real(8), intent(out) :: xs(m)
    integer :: q, it2, kk
    real(8) :: z, p0, p1, p2, dp, dz
    do q = 0, (m+1)/2 - 1
      z = cos(pi*(q + 0.75d0)/(m + 0.5d0))
      do it2 = 1, 100
        p0 = 1
        p1 = z
        do kk = 1, m-1
          p2 = ((2*kk+1)*z*p1 - kk*p0)/(kk+1)
          p0 = p1
          p1 = p2
        end do
        dp = m*(z*p1 - p0)/(z*z - 1)
        dz = p1/dp
        z = z - dz
        if (abs(dz) < 1d-16) exit
      end do
      xs(m-q) = z
      xs(q+1) = -z
    end do
  end subroutine

end program
