!--------------------------------------------------------------------------!
! phsph: Fortran 2008 iso_c_binding interface to libphsph.so, the MI355X  !
! implementation of the SPH hot path.  The wrappers keep the argument     !
! lists of the reference routines they stand in for:                      !
!   phsph_maketree       -> maketree       (src/main/kdtree.F90:117)      !
!   phsph_densityiterate -> densityiterate (src/main/dens.F90:117)        !
!   phsph_cons2prim      -> cons2prim_everything (cons2prim.f90:274)      !
!   phsph_force          -> force          (src/main/force.F90:193)       !
!   phsph_step_*, phsph_substep_sph -> the particle loops of step         !
!                          (step_leapfrog.f90:150-696, substepping.F90:241)!
! Implicit module inputs of the reference (massoftype, hfact, tolh, the   !
! shock-capturing and timestep options, the box) travel in phsph_params, !
! set once with phsph_init.  A non-zero status from the library is passed !
! to the caller's error handler (in Phantom: fatal()).                    !
!--------------------------------------------------------------------------!
module phsph
 use, intrinsic :: iso_c_binding
 implicit none
 private

 integer, parameter, public :: PHSPH_KERNEL_CUBIC = 0, PHSPH_KERNEL_QUINTIC = 1
 integer, parameter, public :: PHSPH_PREC_FP64 = 0, PHSPH_PREC_MIXED = 1

 ! layout of phsph_params (include/phsph.h)
 type, bind(c), public :: phsph_params
    integer(c_int32_t) :: kernel, periodic, maxvxyzu, mhd, use_dustfrac, ndustsmall
    integer(c_int32_t) :: gravity, ind_timesteps, ieos, maxp, precision, idrag
    integer(c_int32_t) :: ishock_heating, ipdv_heating, iresistive_heating, ilimitdustflux
    real(c_double)     :: xmin, xmax, ymin, ymax, zmin, zmax
    real(c_double)     :: hfact, tolh, massoftype_gas, gamma, polyk
    real(c_double)     :: alpha, alphamax, alphau, alphaB, beta
    real(c_double)     :: C_cour, C_force, dtmax, psidecayfac, overcleanfac, K_code
    real(c_double)     :: tree_accuracy
    real(c_double)     :: grainsize, graindens, seff
 end type phsph_params

 ! type(kdnode) of a non-GRAVITY build (dtype_kdtree.F90:53-68)
 type, bind(c), public :: phsph_kdnode
    real(c_double)     :: xcen(3), size, hmax, mass
    integer(c_int32_t) :: leftchild, rightchild, parent, idum
 end type phsph_kdnode

 ! phsph_step_state, phsph_step_clock, phsph_step_summary (include/phsph.h):
 ! whole steps on device-resident arrays (c_ptr of HBM allocations)
 type, bind(c), public :: phsph_step_state_t
    type(c_ptr) :: xyzh, vxyzu, fxyzu, vpred, Bevol, dBevol, Bpred
    type(c_ptr) :: dustevol, ddustevol, dustpred, dustfrac, deltav, tstop
    type(c_ptr) :: divcurlv, divcurlB, divBsymm, alphaind, gradh, dvdx, eos_vars, poten
    type(c_ptr) :: iphase, ibin, ibin_old, ibin_wake, twas
    integer(c_int32_t) :: np, ldxyzh, ldvxyzu, ldfxyzu, lddivcurlv, ldgradh, ldeos, pad
 end type phsph_step_state_t

 type, bind(c), public :: phsph_step_clock_t
    real(c_double)     :: time, tlast, dt, dtcourant, dtforce
    integer(c_int32_t) :: nbinmax, istepfrac, ibinnow, pad
 end type phsph_step_clock_t

 type, bind(c), public :: phsph_step_summary_t
    real(c_double)     :: dt, dtnew, dterr
    integer(c_int32_t) :: its, nactive, nwake, nvfloorp, nvfloorps, nvfloorc, ngraph, ninst
 end type phsph_step_summary_t

 type(c_ptr), save :: ctx = c_null_ptr

 public :: phsph_default, phsph_init, phsph_finalise, phsph_errmsg
 public :: phsph_maketree, phsph_densityiterate, phsph_cons2prim, phsph_force
 public :: phsph_neighbour_stats
 public :: phsph_set_bins, phsph_get_nbinmax, phsph_set_active_particles, phsph_set_poten
 public :: phsph_init_step, phsph_step_predictor, phsph_substep_sph, phsph_step_predict_sph
 public :: phsph_step_corrector, phsph_step_reset, phsph_step_graph

 interface
    subroutine c_default_params(p, kernel) bind(c, name='phsph_default_params')
       import :: phsph_params, c_int32_t
       type(phsph_params), intent(out) :: p
       integer(c_int32_t), value :: kernel
    end subroutine
    integer(c_int) function c_create(p, device, ctx) bind(c, name='phsph_create')
       import :: phsph_params, c_int, c_int32_t, c_ptr
       type(phsph_params), intent(in) :: p
       integer(c_int32_t), value :: device
       type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function c_destroy(ctx) bind(c, name='phsph_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
    end function
    type(c_ptr) function c_last_error(ctx) bind(c, name='phsph_last_error')
       import :: c_ptr
       type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_maketree(ctx, xyzh, ld, np, iphase, leaf, ncells) bind(c, name='phsph_maketree')
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       type(c_ptr), value :: ctx, xyzh, iphase, leaf
       integer(c_int32_t), value :: ld, np
       integer(c_int64_t), intent(out) :: ncells
    end function
    integer(c_int) function c_densityiterate(ctx, icall, np, xyzh, ldx, vxyzu, ldv, fxyzu, ldf, fext, Bevol, &
                                             iphase, divcurlv, lddcv, divcurlB, alphaind, gradh, ldgradh, dvdx) &
                                             bind(c, name='phsph_densityiterate')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx, xyzh, vxyzu, fxyzu, fext, Bevol, iphase, divcurlv, divcurlB, alphaind, gradh, dvdx
       integer(c_int32_t), value :: icall, np, ldx, ldv, ldf, lddcv, ldgradh
    end function
    integer(c_int) function c_cons2prim(ctx, np, xyzh, ldx, vxyzu, ldv, dvdx, alphaind, dustevol, dustfrac, &
                                        eos_vars, ldeos, iphase) bind(c, name='phsph_cons2prim')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx, xyzh, vxyzu, dvdx, alphaind, dustevol, dustfrac, eos_vars, iphase
       integer(c_int32_t), value :: np, ldx, ldv, ldeos
    end function
    integer(c_int) function c_force(ctx, icall, np, xyzh, ldx, vxyzu, ldv, fxyzu, ldf, divcurlv, lddcv, divcurlB, &
                                    Bevol, dBevol, divBsymm, dustfrac, ddustevol, deltav, tstop, gradh, ldgradh, &
                                    alphaind, eos_vars, ldeos, iphase, dtcourant, dtforce) bind(c, name='phsph_force')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx, xyzh, vxyzu, fxyzu, divcurlv, divcurlB, Bevol, dBevol, divBsymm, dustfrac, &
                             ddustevol, deltav, tstop, gradh, alphaind, eos_vars, iphase
       integer(c_int32_t), value :: icall, np, ldx, ldv, ldf, lddcv, ldgradh, ldeos
       real(c_double), intent(out) :: dtcourant, dtforce
    end function
    integer(c_int) function c_neighbour_stats(ctx, trialmean, actualmean, maxtrial, maxactual, nrhocalc, &
                                              nactualtot) bind(c, name='phsph_get_neighbour_stats')
       import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: trialmean, actualmean
       integer(c_int32_t), intent(out) :: maxtrial, maxactual
       integer(c_int64_t), intent(out) :: nrhocalc, nactualtot
    end function
    integer(c_int) function c_set_bins(ctx, ibin, ibin_old, ibin_wake, nbinmax, istepfrac, ibinnow) &
                                       bind(c, name='phsph_set_bins')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx, ibin, ibin_old, ibin_wake
       integer(c_int32_t), value :: nbinmax, istepfrac, ibinnow
    end function
    integer(c_int) function c_set_poten(ctx, poten) bind(c, name='phsph_set_poten')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, poten
    end function
    integer(c_int) function c_get_nbinmax(ctx, nbinmax) bind(c, name='phsph_get_nbinmax')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: nbinmax
    end function
    integer(c_int) function c_set_active(ctx, np, xyzh, ldx, iphase, ibin, nbinmax, istepfrac, nactive, nalive, &
                                         ibinnow) bind(c, name='phsph_set_active_particles')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx, xyzh, iphase, ibin
       integer(c_int32_t), value :: np, ldx, nbinmax, istepfrac
       integer(c_int32_t), intent(out) :: nactive, nalive, ibinnow
    end function
    integer(c_int) function c_init_step(ctx, np, time, ibin, nbinmax, twas) bind(c, name='phsph_init_step')
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx, ibin, twas
       integer(c_int32_t), value :: np, nbinmax
       real(c_double), value :: time
    end function
    integer(c_int) function c_step_predictor(ctx, np, t, dtsph, xyzh, ldx, vxyzu, ldv, fxyzu, ldf, Bevol, dBevol, &
                                             dustevol, ddustevol, iphase, ibin, ibin_old, twas, ufloor, nvfloor) &
                                             bind(c, name='phsph_step_predictor')
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx, xyzh, vxyzu, fxyzu, Bevol, dBevol, dustevol, ddustevol, iphase, ibin, ibin_old, twas
       integer(c_int32_t), value :: np, ldx, ldv, ldf
       real(c_double), value :: t, dtsph, ufloor
       integer(c_int32_t), intent(out) :: nvfloor
    end function
    integer(c_int) function c_substep_sph(ctx, np, dt, xyzh, ldx, vxyzu, ldv) bind(c, name='phsph_substep_sph')
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx, xyzh, vxyzu
       integer(c_int32_t), value :: np, ldx, ldv
       real(c_double), value :: dt
    end function
    integer(c_int) function c_step_predict_sph(ctx, np, timei, dtsph, xyzh, ldx, vxyzu, ldv, fxyzu, ldf, vpred, &
                                               Bevol, dBevol, Bpred, dustevol, ddustevol, dustpred, divcurlv, &
                                               lddcv, eos_vars, ldeos, alphaind, iphase, twas, ufloor, nvfloor) &
                                               bind(c, name='phsph_step_predict_sph')
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx, xyzh, vxyzu, fxyzu, vpred, Bevol, dBevol, Bpred, dustevol, ddustevol, dustpred
       type(c_ptr), value :: divcurlv, eos_vars, alphaind, iphase, twas
       integer(c_int32_t), value :: np, ldx, ldv, ldf, lddcv, ldeos
       real(c_double), value :: timei, dtsph, ufloor
       integer(c_int32_t), intent(out) :: nvfloor
    end function
    integer(c_int) function c_step_corrector(ctx, np, timei, dtsph, xyzh, ldx, vxyzu, ldv, fxyzu, ldf, vpred, &
                                             Bevol, dBevol, dustevol, ddustevol, iphase, ibin, ibin_wake, nbinmax, &
                                             twas, ufloor, errmax, v2sum, npc, nwake, nvfloor) &
                                             bind(c, name='phsph_step_corrector')
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx, xyzh, vxyzu, fxyzu, vpred, Bevol, dBevol, dustevol, ddustevol, iphase
       type(c_ptr), value :: ibin, ibin_wake, twas
       integer(c_int32_t), value :: np, ldx, ldv, ldf, nbinmax
       real(c_double), value :: timei, dtsph, ufloor
       real(c_double), intent(out) :: errmax, v2sum
       integer(c_int32_t), intent(out) :: npc, nwake, nvfloor
    end function
    integer(c_int) function c_step_graph(ctx, st, clk, nsub, ufloor, tolv, maxits, graphs, summaries) &
                                         bind(c, name='phsph_step_graph')
       import :: c_int, c_int32_t, c_double, c_ptr, phsph_step_state_t, phsph_step_clock_t, phsph_step_summary_t
       type(c_ptr), value :: ctx
       type(phsph_step_state_t), intent(in) :: st
       type(phsph_step_clock_t), intent(inout) :: clk
       integer(c_int32_t), value :: nsub, maxits, graphs
       real(c_double), value :: ufloor, tolv
       type(phsph_step_summary_t), intent(out) :: summaries(*)
    end function
    integer(c_int) function c_step_reset(ctx, np, dtsph, vxyzu, ldv, fxyzu, ldf, vpred, Bevol, dBevol, Bpred, &
                                         dustevol, ddustevol, dustpred, iphase) bind(c, name='phsph_step_reset')
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx, vxyzu, fxyzu, vpred, Bevol, dBevol, Bpred, dustevol, ddustevol, dustpred, iphase
       integer(c_int32_t), value :: np, ldv, ldf
       real(c_double), value :: dtsph
    end function
 end interface

contains

!-- c_loc of an array, or NULL for a zero-size (switched-off) one
type(c_ptr) function ptr_or_null(a) result(pp)
 real(c_double), intent(in), target :: a(..)
 pp = c_null_ptr
 if (size(a) > 0) pp = c_loc(a)
end function ptr_or_null

type(c_ptr) function ptr_or_null_i1(a) result(pp)
 integer(c_int8_t), intent(in), target :: a(:)
 pp = c_null_ptr
 if (size(a) > 0) pp = c_loc(a)
end function ptr_or_null_i1

!-- reference defaults for a kernel (options.f90, shock_capturing.f90, timestep.f90)
function phsph_default(kernel) result(p)
 integer, intent(in) :: kernel
 type(phsph_params)  :: p
 call c_default_params(p, int(kernel, c_int32_t))
end function phsph_default

integer function phsph_init(p, device) result(ierr)
 type(phsph_params), intent(in) :: p
 integer, intent(in), optional  :: device
 integer(c_int32_t) :: dev
 dev = 0
 if (present(device)) dev = int(device, c_int32_t)
 ierr = c_create(p, dev, ctx)
end function phsph_init

subroutine phsph_finalise()
 integer(c_int) :: ierr
 if (c_associated(ctx)) ierr = c_destroy(ctx)
 ctx = c_null_ptr
end subroutine phsph_finalise

function phsph_errmsg() result(msg)
 character(len=256) :: msg
 character(kind=c_char), pointer :: s(:)
 type(c_ptr) :: p
 integer :: i
 msg = ' '
 p = c_last_error(ctx)
 if (.not.c_associated(p)) return
 call c_f_pointer(p, s, [256])
 do i=1,256
    if (s(i) == c_null_char) exit
    msg(i:i) = s(i)
 enddo
end function phsph_errmsg

!-- maketree(node,xyzh,np,leaf_is_active,ncells) without the node array,
!   which stays on the device (phsph_get_tree mirrors it when needed)
integer function phsph_maketree(xyzh, np, iphase, leaf_is_active, ncells) result(ierr)
 real(c_double),     intent(inout), target :: xyzh(:,:)
 integer,            intent(in)            :: np
 integer(c_int8_t),  intent(in),    target :: iphase(:)
 integer(c_int32_t), intent(out),   target :: leaf_is_active(:)
 integer(kind=8),    intent(out)           :: ncells
 integer(c_int64_t) :: nc
 ierr = c_maketree(ctx, c_loc(xyzh), int(size(xyzh,1), c_int32_t), int(np, c_int32_t), c_loc(iphase), &
                   c_loc(leaf_is_active), nc)
 ncells = nc
end function phsph_maketree

!-- densityiterate(icall,npart,nactive,xyzh,vxyzu,divcurlv,divcurlB,Bevol,
!   stressmax,fxyzu,fext,alphaind,gradh,rad,radprop,dvdx) for hydro/MHD
integer function phsph_densityiterate(icall, npart, xyzh, vxyzu, divcurlv, divcurlB, Bevol, fxyzu, fext, &
                                      alphaind, gradh, dvdx, iphase) result(ierr)
 integer,           intent(in)            :: icall, npart
 real(c_double),    intent(inout), target :: xyzh(:,:)
 real(c_double),    intent(in),    target :: vxyzu(:,:), fxyzu(:,:), fext(:,:), Bevol(:,:)
 real(c_float),     intent(inout), target :: divcurlv(:,:), divcurlB(:,:), alphaind(:,:), gradh(:,:), dvdx(:,:)
 integer(c_int8_t), intent(in),    target :: iphase(:)
 type(c_ptr) :: pB, pdB
 pB = c_null_ptr
 pdB = c_null_ptr
 if (size(Bevol) > 0) pB = c_loc(Bevol)
 if (size(divcurlB) > 0) pdB = c_loc(divcurlB)
 ierr = c_densityiterate(ctx, int(icall, c_int32_t), int(npart, c_int32_t), c_loc(xyzh), &
                         int(size(xyzh,1), c_int32_t), c_loc(vxyzu), int(size(vxyzu,1), c_int32_t), &
                         c_loc(fxyzu), int(size(fxyzu,1), c_int32_t), c_loc(fext), pB, c_loc(iphase), &
                         c_loc(divcurlv), int(size(divcurlv,1), c_int32_t), pdB, c_loc(alphaind), &
                         c_loc(gradh), int(size(gradh,1), c_int32_t), c_loc(dvdx))
end function phsph_densityiterate

!-- cons2prim_everything for ieos 1 and 2 (eos_vars(1:2,:) = P, c_s)
integer function phsph_cons2prim(npart, xyzh, vxyzu, dvdx, alphaind, eos_vars) result(ierr)
 integer,        intent(in)            :: npart
 real(c_double), intent(in),    target :: xyzh(:,:), vxyzu(:,:)
 real(c_float),  intent(in),    target :: dvdx(:,:)
 real(c_float),  intent(inout), target :: alphaind(:,:)
 real(c_double), intent(inout), target :: eos_vars(:,:)
 ierr = c_cons2prim(ctx, int(npart, c_int32_t), c_loc(xyzh), int(size(xyzh,1), c_int32_t), c_loc(vxyzu), &
                    int(size(vxyzu,1), c_int32_t), c_loc(dvdx), c_loc(alphaind), c_null_ptr, c_null_ptr, &
                    c_loc(eos_vars), int(size(eos_vars,1), c_int32_t), c_null_ptr)
end function phsph_cons2prim

!-- force(icall,npart,xyzh,vxyzu,fxyzu,divcurlv,divcurlB,Bevol,dBevol,...)
!   for hydro (+ MHD when the arrays are non-empty); global timestep
integer function phsph_force(icall, npart, xyzh, vxyzu, fxyzu, divcurlv, divcurlB, Bevol, dBevol, divBsymm, &
                             gradh, alphaind, eos_vars, iphase, dtcourant, dtforce) result(ierr)
 integer,           intent(in)            :: icall, npart
 real(c_double),    intent(in),    target :: xyzh(:,:), vxyzu(:,:), Bevol(:,:), eos_vars(:,:)
 real(c_double),    intent(inout), target :: fxyzu(:,:), dBevol(:,:)
 real(c_float),     intent(inout), target :: divcurlv(:,:), divBsymm(:)
 real(c_float),     intent(in),    target :: divcurlB(:,:), gradh(:,:), alphaind(:,:)
 integer(c_int8_t), intent(in),    target :: iphase(:)
 real(c_double),    intent(out)           :: dtcourant, dtforce
 type(c_ptr) :: pB, pdB, pdBe, pdBs
 pB = c_null_ptr; pdB = c_null_ptr; pdBe = c_null_ptr; pdBs = c_null_ptr
 if (size(Bevol) > 0) then
    pB = c_loc(Bevol); pdB = c_loc(divcurlB); pdBe = c_loc(dBevol); pdBs = c_loc(divBsymm)
 endif
 ierr = c_force(ctx, int(icall, c_int32_t), int(npart, c_int32_t), c_loc(xyzh), int(size(xyzh,1), c_int32_t), &
                c_loc(vxyzu), int(size(vxyzu,1), c_int32_t), c_loc(fxyzu), int(size(fxyzu,1), c_int32_t), &
                c_loc(divcurlv), int(size(divcurlv,1), c_int32_t), pdB, pB, pdBe, pdBs, c_null_ptr, c_null_ptr, &
                c_null_ptr, c_null_ptr, c_loc(gradh), int(size(gradh,1), c_int32_t), c_loc(alphaind), &
                c_loc(eos_vars), int(size(eos_vars,1), c_int32_t), c_loc(iphase), dtcourant, dtforce)
end function phsph_force

!-- get_neighbour_stats (dens.F90:1137)
subroutine phsph_neighbour_stats(trialmean, actualmean, maxtrial, maxactual, nrhocalc, nactualtot)
 real,            intent(out) :: trialmean, actualmean
 integer,         intent(out) :: maxtrial, maxactual
 integer(kind=8), intent(out) :: nrhocalc, nactualtot
 real(c_double)     :: tm, am
 integer(c_int32_t) :: mt, ma
 integer(c_int64_t) :: nr, na
 integer(c_int)     :: ierr
 ierr = c_neighbour_stats(ctx, tm, am, mt, ma, nr, na)
 trialmean = tm; actualmean = am; maxtrial = mt; maxactual = ma; nrhocalc = nr; nactualtot = na
end subroutine phsph_neighbour_stats

!-- block timesteps: register part:ibin, ibin_old, ibin_wake and the
!   timestep_ind state (nbinmax, istepfrac, ibinnow) for the next force calls
integer function phsph_set_bins(ibin, ibin_old, ibin_wake, nbinmax, istepfrac, ibinnow) result(ierr)
 integer(c_int8_t), intent(inout), target :: ibin(:), ibin_wake(:)
 integer(c_int8_t), intent(in),    target :: ibin_old(:)
 integer,           intent(in)            :: nbinmax, istepfrac, ibinnow
 ierr = c_set_bins(ctx, c_loc(ibin), c_loc(ibin_old), c_loc(ibin_wake), int(nbinmax, c_int32_t), &
                   int(istepfrac, c_int32_t), int(ibinnow, c_int32_t))
end function phsph_set_bins

!-- part:poten, written by force with self-gravity (force.F90:2927)
integer function phsph_set_poten(poten) result(ierr)
 real(c_float), intent(inout), target :: poten(:)
 ierr = c_set_poten(ctx, c_loc(poten))
end function phsph_set_poten

!-- nbinmax after the last force call (force.F90:792)
integer function phsph_get_nbinmax(nbinmax) result(ierr)
 integer(kind=1), intent(out) :: nbinmax
 integer(c_int32_t) :: nb
 ierr = c_get_nbinmax(ctx, nb)
 nbinmax = int(nb, kind=1)
end function phsph_get_nbinmax

!-- set_active_particles (utils_indtimesteps.f90:114)
integer function phsph_set_active_particles(npart, xyzh, iphase, ibin, nbinmax, istepfrac, nactive, nalive, &
                                            ibinnow) result(ierr)
 integer,           intent(in)            :: npart, nbinmax, istepfrac
 real(c_double),    intent(in),    target :: xyzh(:,:)
 integer(c_int8_t), intent(inout), target :: iphase(:)
 integer(c_int8_t), intent(in),    target :: ibin(:)
 integer,           intent(out)           :: nactive, nalive
 integer(kind=1),   intent(out)           :: ibinnow
 integer(c_int32_t) :: na, nl, bn
 ierr = c_set_active(ctx, int(npart, c_int32_t), c_loc(xyzh), int(size(xyzh,1), c_int32_t), c_loc(iphase), &
                     c_loc(ibin), int(nbinmax, c_int32_t), int(istepfrac, c_int32_t), na, nl, bn)
 nactive = na; nalive = nl; ibinnow = int(bn, kind=1)
end function phsph_set_active_particles

!-- init_step (step_leapfrog.f90:44): twas = time + dt(ibin)/2 (block timesteps)
integer function phsph_init_step(npart, time, ibin, nbinmax, twas) result(ierr)
 integer,           intent(in)            :: npart, nbinmax
 real(c_double),    intent(in)            :: time
 integer(c_int8_t), intent(inout), target :: ibin(:)
 real(c_double),    intent(inout), target :: twas(:)
 ierr = c_init_step(ctx, int(npart, c_int32_t), time, c_loc(ibin), int(nbinmax, c_int32_t), c_loc(twas))
end function phsph_init_step

!-- predictor loop of step (step_leapfrog.f90:150-237).  Zero-size arrays
!   stand for physics that is off (Bevol/dBevol for MHD, dustevol/ddustevol
!   for dust, ibin/ibin_old/twas for block timesteps)
integer function phsph_step_predictor(npart, t, dtsph, xyzh, vxyzu, fxyzu, Bevol, dBevol, dustevol, ddustevol, &
                                      iphase, ibin, ibin_old, twas, ufloor, nvfloor) result(ierr)
 integer,           intent(in)            :: npart
 real(c_double),    intent(in)            :: t, dtsph, ufloor
 real(c_double),    intent(in),    target :: xyzh(:,:), fxyzu(:,:), dBevol(:,:), ddustevol(:,:), twas(:)
 real(c_double),    intent(inout), target :: vxyzu(:,:), Bevol(:,:), dustevol(:,:)
 integer(c_int8_t), intent(in),    target :: iphase(:), ibin(:)
 integer(c_int8_t), intent(inout), target :: ibin_old(:)
 integer,           intent(out)           :: nvfloor
 integer(c_int32_t) :: nf
 ierr = c_step_predictor(ctx, int(npart, c_int32_t), t, dtsph, c_loc(xyzh), int(size(xyzh,1), c_int32_t), &
                         c_loc(vxyzu), int(size(vxyzu,1), c_int32_t), c_loc(fxyzu), int(size(fxyzu,1), c_int32_t), &
                         ptr_or_null(Bevol), ptr_or_null(dBevol), ptr_or_null(dustevol), ptr_or_null(ddustevol), &
                         ptr_or_null_i1(iphase), ptr_or_null_i1(ibin), ptr_or_null_i1(ibin_old), ptr_or_null(twas), &
                         ufloor, nf)
 nvfloor = nf
end function phsph_step_predictor

!-- substep_sph (substepping.F90:241): x += dt*v
integer function phsph_substep_sph(dt, npart, xyzh, vxyzu) result(ierr)
 real(c_double), intent(in)            :: dt
 integer,        intent(in)            :: npart
 real(c_double), intent(inout), target :: xyzh(:,:)
 real(c_double), intent(in),    target :: vxyzu(:,:)
 ierr = c_substep_sph(ctx, int(npart, c_int32_t), dt, c_loc(xyzh), int(size(xyzh,1), c_int32_t), c_loc(vxyzu), &
                      int(size(vxyzu,1), c_int32_t))
end function phsph_substep_sph

!-- predict_sph loop of step (step_leapfrog.f90:307-400)
integer function phsph_step_predict_sph(npart, timei, dtsph, xyzh, vxyzu, fxyzu, vpred, Bevol, dBevol, Bpred, &
                                        dustevol, ddustevol, dustpred, divcurlv, eos_vars, alphaind, iphase, &
                                        twas, ufloor, nvfloor) result(ierr)
 integer,           intent(in)            :: npart
 real(c_double),    intent(in)            :: timei, dtsph, ufloor
 real(c_double),    intent(inout), target :: xyzh(:,:), vpred(:,:), Bpred(:,:), dustpred(:,:)
 real(c_double),    intent(in),    target :: vxyzu(:,:), fxyzu(:,:), Bevol(:,:), dBevol(:,:), dustevol(:,:)
 real(c_double),    intent(in),    target :: ddustevol(:,:), eos_vars(:,:), twas(:)
 real(c_float),     intent(in),    target :: divcurlv(:,:)
 real(c_float),     intent(inout), target :: alphaind(:,:)
 integer(c_int8_t), intent(in),    target :: iphase(:)
 integer,           intent(out)           :: nvfloor
 integer(c_int32_t) :: nf
 ierr = c_step_predict_sph(ctx, int(npart, c_int32_t), timei, dtsph, c_loc(xyzh), int(size(xyzh,1), c_int32_t), &
                           c_loc(vxyzu), int(size(vxyzu,1), c_int32_t), c_loc(fxyzu), int(size(fxyzu,1), c_int32_t), &
                           c_loc(vpred), ptr_or_null(Bevol), ptr_or_null(dBevol), ptr_or_null(Bpred), &
                           ptr_or_null(dustevol), ptr_or_null(ddustevol), ptr_or_null(dustpred), c_loc(divcurlv), &
                           int(size(divcurlv,1), c_int32_t), c_loc(eos_vars), int(size(eos_vars,1), c_int32_t), &
                           c_loc(alphaind), ptr_or_null_i1(iphase), ptr_or_null(twas), ufloor, nf)
 nvfloor = nf
end function phsph_step_predict_sph

!-- one pass of the corrector loop (step_leapfrog.f90:482-625); errmax,
!   v2sum and npc are the inputs of check_velocity_error
integer function phsph_step_corrector(npart, timei, dtsph, xyzh, vxyzu, fxyzu, vpred, Bevol, dBevol, dustevol, &
                                      ddustevol, iphase, ibin, ibin_wake, nbinmax, twas, ufloor, errmax, v2sum, &
                                      npc, nwake, nvfloor) result(ierr)
 integer,           intent(in)            :: npart, nbinmax
 real(c_double),    intent(in)            :: timei, dtsph, ufloor
 real(c_double),    intent(in),    target :: xyzh(:,:), fxyzu(:,:), vpred(:,:), dBevol(:,:), ddustevol(:,:)
 real(c_double),    intent(inout), target :: vxyzu(:,:), Bevol(:,:), dustevol(:,:), twas(:)
 integer(c_int8_t), intent(in),    target :: iphase(:)
 integer(c_int8_t), intent(inout), target :: ibin(:), ibin_wake(:)
 real(c_double),    intent(out)           :: errmax, v2sum
 integer,           intent(out)           :: npc, nwake, nvfloor
 integer(c_int32_t) :: n1, n2, n3
 ierr = c_step_corrector(ctx, int(npart, c_int32_t), timei, dtsph, c_loc(xyzh), int(size(xyzh,1), c_int32_t), &
                         c_loc(vxyzu), int(size(vxyzu,1), c_int32_t), c_loc(fxyzu), int(size(fxyzu,1), c_int32_t), &
                         ptr_or_null(vpred), ptr_or_null(Bevol), ptr_or_null(dBevol), ptr_or_null(dustevol), &
                         ptr_or_null(ddustevol), ptr_or_null_i1(iphase), ptr_or_null_i1(ibin), &
                         ptr_or_null_i1(ibin_wake), int(nbinmax, c_int32_t), ptr_or_null(twas), ufloor, errmax, &
                         v2sum, n1, n2, n3)
 npc = n1; nwake = n2; nvfloor = n3
end function phsph_step_corrector

!-- until_converged loop of step (step_leapfrog.f90:650-696)
integer function phsph_step_reset(npart, dtsph, vxyzu, fxyzu, vpred, Bevol, dBevol, Bpred, dustevol, ddustevol, &
                                  dustpred, iphase) result(ierr)
 integer,           intent(in)            :: npart
 real(c_double),    intent(in)            :: dtsph
 real(c_double),    intent(inout), target :: vxyzu(:,:), vpred(:,:), Bevol(:,:), Bpred(:,:), dustevol(:,:)
 real(c_double),    intent(inout), target :: dustpred(:,:)
 real(c_double),    intent(in),    target :: fxyzu(:,:), dBevol(:,:), ddustevol(:,:)
 integer(c_int8_t), intent(in),    target :: iphase(:)
 ierr = c_step_reset(ctx, int(npart, c_int32_t), dtsph, c_loc(vxyzu), int(size(vxyzu,1), c_int32_t), c_loc(fxyzu), &
                     int(size(fxyzu,1), c_int32_t), c_loc(vpred), ptr_or_null(Bevol), ptr_or_null(dBevol), &
                     ptr_or_null(Bpred), ptr_or_null(dustevol), ptr_or_null(ddustevol), ptr_or_null(dustpred), &
                     ptr_or_null_i1(iphase))
end function phsph_step_reset

!-- nsub steps of the evolve loop inside the library on device-resident state
!   (evolve.f90:105-257 with step, step_leapfrog.f90:91-843)
integer function phsph_step_graph(st, clk, nsub, ufloor, tolv, maxits, graphs, summaries) result(ierr)
 type(phsph_step_state_t), intent(in) :: st
 type(phsph_step_clock_t), intent(inout) :: clk
 integer, intent(in) :: nsub, maxits
 real(c_double), intent(in) :: ufloor, tolv
 logical, intent(in) :: graphs
 type(phsph_step_summary_t), intent(out) :: summaries(nsub)
 integer(c_int32_t) :: ig
 ig = 0
 if (graphs) ig = 1
 ierr = c_step_graph(ctx, st, clk, int(nsub, c_int32_t), ufloor, tolv, int(maxits, c_int32_t), ig, summaries)
end function phsph_step_graph

end module phsph
