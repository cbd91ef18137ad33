!--------------------------------------------------------------------------!
! A Fortran host running whole leapfrog steps through module phsph, with   !
! the control flow of Phantom's step (step_leapfrog.f90:91-760): predictor,!
! drift (substep_sph), predict_sph, derivs (maketree -> densityiterate ->  !
! cons2prim -> force), then corrector passes with the velocity-error test  !
! until converged.  The particles are those of the reference's step test   !
! (test_step.F90:63-160): a cubic lattice moving at v = (1,1,1) through a  !
! periodic unit box with zero pressure and no dissipation, 10 steps of     !
! dt = 0.2; h must stay within 3e-4 of hzero and every force must vanish.  !
! Exit status 0 on success, 1 on a library error, 2 on a wrong answer.     !
!--------------------------------------------------------------------------!
program step_driver
 use, intrinsic :: iso_c_binding
 use phsph
 implicit none
 integer, parameter :: nx = 20, np = nx**3, nsteps = 10, maxits = 30
 real(8), parameter :: rhozero = 7.5d0, tolv = 1.d-2
 type(phsph_params) :: p
 real(c_double), allocatable :: xyzh(:,:), vxyzu(:,:), vpred(:,:), fxyzu(:,:), fext(:,:), eos_vars(:,:)
 real(c_double), allocatable :: none2(:,:), none1(:)
 real(c_float),  allocatable :: divcurlv(:,:), divcurlB(:,:), alphaind(:,:), gradh(:,:), dvdx(:,:), divBsymm(:)
 integer(c_int8_t), allocatable :: iphase(:), nobin(:)
 integer(c_int32_t), allocatable :: leaf_is_active(:)
 integer(kind=8) :: ncells
 integer :: i, j, k, n, ierr, istep, its, nfl, npc, nwake
 real(8) :: dx, dt, t, timei, hzero, errmax, v2sum, dtcourant, dtforce, herr, fmax
 logical :: converged

 p = phsph_default(PHSPH_KERNEL_CUBIC)
 p%massoftype_gas = rhozero/np
 p%tolh = 1.d-5
 p%maxp = np
 p%gamma = 5.d0/3.d0
 p%polyk = 0.d0
 p%alpha = 0.; p%alphau = 0.; p%alphaB = 0.
 dt = 2.d0/nsteps
 p%dtmax = dt
 ierr = phsph_init(p)
 if (ierr /= 0) then
    print*, 'phsph_init: ', trim(phsph_errmsg())
    stop 1
 endif
 hzero = p%hfact*(p%massoftype_gas/rhozero)**(1.d0/3.d0)
 allocate(xyzh(4,np), vxyzu(4,np), vpred(4,np), fxyzu(4,np), fext(3,np), eos_vars(3,np))
 allocate(none2(4,0), none1(0), nobin(0))
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
       enddo
    enddo
 enddo
 vxyzu(1:3,:) = 1.d0; vxyzu(4,:) = 0.d0
 fxyzu = 0.d0; fext = 0.d0; alphaind = 0.; iphase = 1_c_int8_t

 ! initial derivatives (initial.F90: derivs before the first step)
 vpred = vxyzu
 call derivs(1)
 t = 0.d0
 do istep=1,nsteps
    ! predictor (step_leapfrog.f90:183-237) and drift (:279)
    ierr = phsph_step_predictor(np, t, dt, xyzh, vxyzu, fxyzu, none2, none2, none2, none2, iphase, nobin, nobin, &
                                none1, 0.d0, nfl)
    if (ierr /= 0) call die('step_predictor')
    ierr = phsph_substep_sph(dt, np, xyzh, vxyzu)
    if (ierr /= 0) call die('substep_sph')
    timei = t + dt
    ! prediction to the full step (:307-400)
    ierr = phsph_step_predict_sph(np, timei, dt, xyzh, vxyzu, fxyzu, vpred, none2, none2, none2, none2, none2, &
                                  none2, divcurlv, eos_vars, alphaind, iphase, none1, 0.d0, nfl)
    if (ierr /= 0) call die('step_predict_sph')
    call derivs(1)
    ! corrector (:482-714)
    its = 0
    converged = .false.
    do while (its < maxits .and. .not.converged)
       its = its + 1
       ierr = phsph_step_corrector(np, timei, dt, xyzh, vxyzu, fxyzu, vpred, none2, none2, none2, none2, iphase, &
                                   nobin, nobin, 0, none1, 0.d0, errmax, v2sum, npc, nwake, nfl)
       if (ierr /= 0) call die('step_corrector')
       ! check_velocity_error (:769-843)
       if (npc > 0 .and. v2sum/npc > tiny(v2sum)) then
          errmax = errmax/sqrt(v2sum/npc)
       else
          errmax = 0.d0
       endif
       converged = (errmax < tolv)
       if (.not.converged) then
          ierr = phsph_step_reset(np, dt, vxyzu, fxyzu, vpred, none2, none2, none2, none2, none2, none2, iphase)
          if (ierr /= 0) call die('step_reset')
          call derivs(2)
       endif
    enddo
    t = t + dt
    herr = maxval(abs(xyzh(4,:) - hzero))/hzero
    fmax = maxval(abs(fxyzu))
    print "(a,i3,a,es10.3,a,es10.3,a,i2)", ' step ', istep, ' max|h/hzero-1| = ', herr, ' max|f| = ', fmax, &
          ' its = ', its
    if (herr > 3.e-4 .or. fmax > tiny(fmax) .or. its /= 1) then
       call phsph_finalise()
       stop 2
    endif
 enddo
 call phsph_finalise()
 print*, 'step_driver: ok'

contains

! derivs (deriv.f90:1-197) on the predicted velocities
subroutine derivs(icall)
 integer, intent(in) :: icall
 if (icall == 1) then
    ierr = phsph_maketree(xyzh, np, iphase, leaf_is_active, ncells)
    if (ierr /= 0) call die('maketree')
    ierr = phsph_densityiterate(1, np, xyzh, vpred, divcurlv, divcurlB, none2, fxyzu, fext, alphaind, gradh, &
                                dvdx, iphase)
    if (ierr /= 0) call die('densityiterate')
 endif
 ierr = phsph_cons2prim(np, xyzh, vpred, dvdx, alphaind, eos_vars)
 if (ierr /= 0) call die('cons2prim')
 ierr = phsph_force(icall, np, xyzh, vpred, fxyzu, divcurlv, divcurlB, none2, none2, divBsymm, gradh, alphaind, &
                    eos_vars, iphase, dtcourant, dtforce)
 if (ierr /= 0) call die('force')
end subroutine derivs

subroutine die(where)
 character(len=*), intent(in) :: where
 print*, where, ': ', trim(phsph_errmsg())
 call phsph_finalise()
 stop 1
end subroutine die

end program step_driver
