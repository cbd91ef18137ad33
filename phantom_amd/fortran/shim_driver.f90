!--------------------------------------------------------------------------!
! A Fortran host calling the MI355X hot path through module phsph, the way !
! Phantom's deriv.f90 calls build_tree -> densityiterate -> cons2prim ->   !
! force.  It sets up the derivative unit test's lattice (test_derivs.f90:  !
! 128-160: cubic lattice in the unit box, rho = 5, hfact 1.2, sine         !
! velocity and thermal energy fields) at 24^3 and prints the neighbour     !
! statistics, which on an exact lattice are 57 per particle with exactly   !
! two density evaluations each (test_derivs.f90:204-212).                  !
! Exit status 0 on success, 1 on a library error, 2 on a wrong answer.     !
!--------------------------------------------------------------------------!
program shim_driver
 use, intrinsic :: iso_c_binding
 use phsph
 implicit none
 integer, parameter :: nx = 24, np = nx**3
 real(8), parameter :: pi = 3.1415926536d0, rhozero = 5.d0
 type(phsph_params) :: p
 real(c_double), allocatable :: xyzh(:,:), vxyzu(:,:), fxyzu(:,:), fext(:,:), Bevol(:,:), eos_vars(:,:), dBevol(:,:)
 real(c_float),  allocatable :: divcurlv(:,:), divcurlB(:,:), alphaind(:,:), gradh(:,:), dvdx(:,:), divBsymm(:)
 integer(c_int8_t), allocatable :: iphase(:)
 integer(c_int32_t), allocatable :: leaf_is_active(:)
 integer(kind=8) :: ncells, nrhocalc, nactualtot
 integer :: i, j, k, n, ierr, maxtrial, maxactual
 real :: trialmean, actualmean
 real(8) :: dx, dtcourant, dtforce

 p = phsph_default(PHSPH_KERNEL_CUBIC)
 p%massoftype_gas = rhozero/np
 p%tolh = 1.d-5
 p%maxp = np
 p%alpha = 0.; p%alphau = 0.; p%alphaB = 0.; p%beta = 0.
 ierr = phsph_init(p)
 if (ierr /= 0) then
    print*, 'phsph_init: ', trim(phsph_errmsg())
    stop 1
 endif
 allocate(xyzh(4,np), vxyzu(4,np), fxyzu(4,np), fext(3,np), Bevol(4,0), eos_vars(3,np), dBevol(4,0))
 allocate(divcurlv(1,np), divcurlB(4,0), alphaind(3,np), gradh(1,np), dvdx(9,np), divBsymm(0))
 allocate(iphase(np), leaf_is_active(2*np+1))
 dx = 1.d0/nx
 n = 0
 do k=1,nx
    do j=1,nx
       do i=1,nx
          n = n + 1
          xyzh(1,n) = -0.5d0 + (i-0.5d0)*dx
          xyzh(2,n) = -0.5d0 + (j-0.5d0)*dx
          xyzh(3,n) = -0.5d0 + (k-0.5d0)*dx
          xyzh(4,n) = p%hfact*dx
          vxyzu(1,n) = 0.5d0/pi*sin(2.d0*pi*(xyzh(1,n)+0.5d0))
          vxyzu(2,n) = 0.5d0/pi*sin(2.d0*pi*(xyzh(1,n)+0.5d0)) - 0.5d0/pi*sin(2.d0*pi*(xyzh(3,n)+0.5d0))
          vxyzu(3,n) = 0.05d0/pi*cos(4.d0*pi*(xyzh(2,n)+0.5d0))
          vxyzu(4,n) = 0.5d0/pi*(3.d0 + sin(2.d0*pi*(xyzh(1,n)+0.5d0)) + cos(2.d0*pi*(xyzh(2,n)+0.5d0)) &
                       + sin(2.d0*pi*(xyzh(3,n)+0.5d0)))
       enddo
    enddo
 enddo
 fxyzu = 0.d0; fext = 0.d0; alphaind = 0.; iphase = 1_c_int8_t

 ierr = phsph_maketree(xyzh, np, iphase, leaf_is_active, ncells)
 if (ierr /= 0) call die('maketree')
 ierr = phsph_densityiterate(1, np, xyzh, vxyzu, divcurlv, divcurlB, Bevol, fxyzu, fext, alphaind, gradh, dvdx, iphase)
 if (ierr /= 0) call die('densityiterate')
 ierr = phsph_cons2prim(np, xyzh, vxyzu, dvdx, alphaind, eos_vars)
 if (ierr /= 0) call die('cons2prim')
 ierr = phsph_force(1, np, xyzh, vxyzu, fxyzu, divcurlv, divcurlB, Bevol, dBevol, divBsymm, gradh, alphaind, &
                    eos_vars, iphase, dtcourant, dtforce)
 if (ierr /= 0) call die('force')
 call phsph_neighbour_stats(trialmean, actualmean, maxtrial, maxactual, nrhocalc, nactualtot)
 print "(a,i8,a,f8.3,a,i4,a,i10,a,f9.6)", ' np=', np, ' mean nneigh=', actualmean, ' max=', maxactual, &
       ' nrhocalc=', nrhocalc, ' gradh(1)=', gradh(1,1)
 print "(a,es12.5,a,es12.5)", ' dtcourant=', dtcourant, ' dtforce=', dtforce
 call phsph_finalise()
 if (abs(actualmean - 57.) > 0. .or. maxactual /= 57 .or. nrhocalc /= 2*np) stop 2
 if (abs(gradh(1,1) - 1.01948) > 1.e-5*1.01948) stop 2
 print*, 'shim_driver: ok'

contains

subroutine die(where)
 character(len=*), intent(in) :: where
 print*, where, ': ', trim(phsph_errmsg())
 call phsph_finalise()
 stop 1
end subroutine die

end program shim_driver
