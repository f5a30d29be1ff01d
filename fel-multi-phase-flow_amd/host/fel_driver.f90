! fel_driver.f90 -- Fortran host for the gfx950 timestep library.
!
! Owns the host arrays in the reference's layout (-2:n+3 padding, FP64) and
! runs the start-up boundary fills and the timestep of the reference
! (srcs/main.f90:303-366, 467-468, 473-748) by calling the library's
! drop-in entry points: the same subroutine names and argument lists as the
! reference, resolved to libfelgpu.so at link time.  Every array is
! registered with the library once, so the whole step stays in device
! memory; host copies are refreshed only on request (fel_drv_download).
!
! Differences from srcs/main.f90, all documented at their call sites:
!  * configuration comes from the caller (fel_cfg) instead of hard-coded
!    assignments (main.f90:91-249);
!  * the slip-length phi argument of bndu/solu_sor4 is slice 1, where the
!    reference passes phi(-2,-2,-2,l) with l = nbub+1 after a loop, one slice
!    past the array (main.f90:337-339,623,625,650);
!  * output cadences (VTK, slices, dumps: main.f90:702-743) are not part of
!    the step.
module fel_driver
  use iso_c_binding
  implicit none
  private

  integer, parameter :: nbub = 1

  ! Run parameters (srcs/main.f90:91-249), filled by the caller.
  type, bind(C) :: fel_cfg
    integer(c_int) :: ni, nj, nk, imon_t
    real(c_double) :: dx, dy, dz, dt
    real(c_double) :: bet_mthinc, rhol, rhog, rmul, rmug, surface_tension, uwall, grvb, grvp
    real(c_double) :: xl, yl, zl, particle_radius
    real(c_double) :: cfl   ! > 0: adaptive dt by caldt each step; 0: fixed dt
  end type fel_cfg

  type(fel_cfg), save :: c
  logical, save :: created = .false.
  ! the modal arrays of the pressure solve in the flow class (FP32 in the
  ! library's FP32 mode: FP32 solve + FP64 refinement) instead of FP64
  logical, save :: modes_flow_class = .false.
  integer, save :: ni, nj, nk, nstep, ipara, ID, ndiv
  integer, save :: nID(6), svall(3), key(8)
  real(8), save :: dxinv, dyinv, dzinv, time, err_div0, err_div, dt
  real(8), save :: center(3), center_pre1, center_pre2, velocity, vel(3), re(3)
  integer, save :: area_pre1, area_pre2, lap

  real(8), allocatable, target, save :: u(:,:,:), v(:,:,:), w(:,:,:), uo(:,:,:), vo(:,:,:), wo(:,:,:)
  real(8), allocatable, target, save :: un(:,:,:), vn(:,:,:), wn(:,:,:), p(:,:,:), po(:,:,:), pn(:,:,:)
  real(8), allocatable, target, save :: phat(:,:,:), dp(:,:,:), div(:,:,:)
  real(8), allocatable, target, save :: rho(:,:,:), rmu(:,:,:), rhon(:,:,:), rmun(:,:,:)
  real(8), allocatable, target, save :: adv_u(:,:,:), adv_v(:,:,:), adv_w(:,:,:)
  real(8), allocatable, target, save :: adv_uo(:,:,:), adv_vo(:,:,:), adv_wo(:,:,:)
  real(8), allocatable, target, save :: prs_u(:,:,:), prs_v(:,:,:), prs_w(:,:,:)
  real(8), allocatable, target, save :: vis_u(:,:,:), vis_v(:,:,:), vis_w(:,:,:)
  real(8), allocatable, target, save :: fst_u(:,:,:), fst_v(:,:,:), fst_w(:,:,:)
  real(8), allocatable, target, save :: fst_un(:,:,:), fst_vn(:,:,:), fst_wn(:,:,:)
  real(8), allocatable, target, save :: sum_fst_u(:,:,:), sum_fst_v(:,:,:), sum_fst_w(:,:,:)
  real(8), allocatable, target, save :: sum_fst_un(:,:,:), sum_fst_vn(:,:,:), sum_fst_wn(:,:,:)
  real(8), allocatable, target, save :: src_u(:,:,:), src_v(:,:,:), src_w(:,:,:)
  real(8), allocatable, target, save :: phix(:,:,:), phiy(:,:,:), phiz(:,:,:)
  real(8), allocatable, target, save :: flphix(:,:,:), flphiy(:,:,:), flphiz(:,:,:)
  real(8), allocatable, target, save :: phi(:,:,:,:), phin(:,:,:,:)
  real(8), allocatable, target, save :: s(:,:,:,:), tau(:,:,:,:)
  real(8), allocatable, target, save :: rn_pp(:,:,:,:), rn_12(:,:,:,:), rn_13(:,:,:,:), rn_23(:,:,:,:)
  real(8), allocatable, target, save :: rnn_pp(:,:,:,:), rnn_12(:,:,:,:), rnn_13(:,:,:,:), rnn_23(:,:,:,:)
  real(8), allocatable, target, save :: rkap(:,:,:,:), rkapn(:,:,:,:)
  real(8), allocatable, target, save :: cf(:,:,:,:)   ! the 45 viscous coefficient arrays
  real(8), allocatable, target, save :: theta_0_array(:,:,:), l1_array(:,:,:), l2_array(:,:,:)
  real(8), allocatable, target, save :: zeta_array(:,:,:), theta_array(:,:,:)
  real(8), allocatable, target, save :: phir_r(:), phir_i(:)
  real(8), allocatable, save :: dyinv_array(:), fftdata(:)
  real(8), allocatable, save :: aw_p(:), ae_p(:), as_p(:), an_p(:), ab_p(:), at_p(:), ap_p(:)
  real(8), allocatable, save :: as_p_fft(:), an_p_fft(:), ap_p_fft(:), atdma(:), btdma_r(:), btdma_i(:)
  real(8), save :: sendjb(1), recvjb(1), sendbuf(1), recvbuf(1)

  public :: fel_cfg

contains

  subroutine alloc3(a)
    real(8), allocatable, intent(inout) :: a(:,:,:)
    allocate(a(-2:ni+3,-2:nj+3,-2:nk+3))
  end subroutine alloc3

  subroutine alloc4(a, m0, m1)
    real(8), allocatable, intent(inout) :: a(:,:,:,:)
    integer, intent(in) :: m0, m1
    allocate(a(-2:ni+3,-2:nj+3,-2:nk+3,m0:m1))
  end subroutine alloc4

  ! Registers an array with the library: device-resident from now on.
  ! upload=1 copies the host contents (fields the caller initialised),
  ! upload=0 starts the device copy at zero.
  subroutine reg(a, nelem, upload)
    real(8), intent(in) :: a(*)
    integer(c_int64_t), intent(in) :: nelem
    integer, intent(in) :: upload
    integer(c_int64_t) :: nbytes
    integer(c_int) :: up
    nbytes = nelem * 8_c_int64_t
    up = upload
    call felgpu_register(a, nbytes, up)
  end subroutine reg

  ! As reg, for the arrays that stay FP64 on the device in the library's
  ! FP32 mode: the volume fraction and its fluxes, the contact-line maps,
  ! the divergence and the pressure correction.
  subroutine reg64(a, nelem, upload)
    real(8), intent(in) :: a(*)
    integer(c_int64_t), intent(in) :: nelem
    integer, intent(in) :: upload
    integer(c_int64_t) :: nbytes
    integer(c_int) :: up
    nbytes = nelem * 8_c_int64_t
    up = upload
    call felgpu_register_fp64(a, nbytes, up)
  end subroutine reg64

  ! Allocates the run's arrays (srcs/allocate.h:7-165, the subset the step
  ! uses); they are zero on the host until the caller stores initial fields.
  subroutine fel_drv_create(cfg) bind(C, name="fel_drv_create")
    type(fel_cfg), intent(in) :: cfg
    integer :: j
    if (created) call fel_drv_destroy()
    c = cfg
    ni = c%ni; nj = c%nj; nk = c%nk
    svall = (/ ni, nj, nk /)
    nID = -1; ipara = 0; ID = 0; ndiv = 1; key = 0
    dxinv = 1.0d0/c%dx; dyinv = 1.0d0/c%dy; dzinv = 1.0d0/c%dz
    call alloc3(u); call alloc3(v); call alloc3(w); call alloc3(uo); call alloc3(vo); call alloc3(wo)
    call alloc3(un); call alloc3(vn); call alloc3(wn); call alloc3(p); call alloc3(po); call alloc3(pn)
    call alloc3(phat); call alloc3(dp); call alloc3(div)
    call alloc3(rho); call alloc3(rmu); call alloc3(rhon); call alloc3(rmun)
    call alloc3(adv_u); call alloc3(adv_v); call alloc3(adv_w)
    call alloc3(adv_uo); call alloc3(adv_vo); call alloc3(adv_wo)
    call alloc3(prs_u); call alloc3(prs_v); call alloc3(prs_w)
    call alloc3(vis_u); call alloc3(vis_v); call alloc3(vis_w)
    call alloc3(fst_u); call alloc3(fst_v); call alloc3(fst_w)
    call alloc3(fst_un); call alloc3(fst_vn); call alloc3(fst_wn)
    call alloc3(sum_fst_u); call alloc3(sum_fst_v); call alloc3(sum_fst_w)
    call alloc3(sum_fst_un); call alloc3(sum_fst_vn); call alloc3(sum_fst_wn)
    call alloc3(src_u); call alloc3(src_v); call alloc3(src_w)
    call alloc3(phix); call alloc3(phiy); call alloc3(phiz)
    call alloc3(flphix); call alloc3(flphiy); call alloc3(flphiz)
    call alloc4(phi, 0, nbub); call alloc4(phin, 0, nbub)
    call alloc4(s, 1, 6); call alloc4(tau, 1, 6)
    call alloc4(rn_pp, 1, 3); call alloc4(rn_12, 1, 3); call alloc4(rn_13, 1, 3); call alloc4(rn_23, 1, 3)
    call alloc4(rnn_pp, 1, 3); call alloc4(rnn_12, 1, 3); call alloc4(rnn_13, 1, 3); call alloc4(rnn_23, 1, 3)
    call alloc4(rkap, 1, 3); call alloc4(rkapn, 1, 3)
    call alloc4(cf, 1, 45)
    call alloc3(theta_0_array); call alloc3(l1_array); call alloc3(l2_array)
    call alloc3(zeta_array); call alloc3(theta_array)
    allocate(phir_r(ni*nj*nk), phir_i(ni*nj*nk))
    allocate(dyinv_array(-2:nj+3), fftdata(2))
    allocate(aw_p(-2:nj+3), ae_p(-2:nj+3), as_p(-2:nj+3), an_p(-2:nj+3), ab_p(-2:nj+3), at_p(-2:nj+3))
    allocate(ap_p(-2:nj+3), as_p_fft(-2:nj+3), an_p_fft(-2:nj+3), ap_p_fft(-2:nj+3))
    allocate(atdma(-2:max(nj,nk)+3), btdma_r(-2:max(nj,nk)+3), btdma_i(-2:max(nj,nk)+3))
    do j = -2, nj+3
      dyinv_array(j) = 1.0d0/c%dy
    end do
    ! fields the caller may leave untouched but that are uploaded (srcs/init.f)
    u = 0; v = 0; w = 0; p = 0; phi = 0
    theta_0_array = 0; l1_array = 0; l2_array = 0; zeta_array = 0
    aw_p = 0; ae_p = 0; as_p = 0; an_p = 0; ab_p = 0; at_p = 0; ap_p = 0
    as_p_fft = 0; an_p_fft = 0; ap_p_fft = 0
    nstep = 0; time = 0.0d0; err_div0 = 0.0d0; err_div = 0.0d0
    center = 0; center_pre1 = 0; center_pre2 = 0; velocity = 0; vel = 0; re = 0
    area_pre1 = 0; area_pre2 = 0; lap = 0
    created = .true.
  end subroutine fel_drv_create

  ! Host address of a field, for the caller to store initial data or read
  ! results: u v w uo vo wo un vn wn p po pn dp div phat rho rmu rhon rmun
  ! phi0 phi1 phin0 phin1 theta_0_array l1_array l2_array zeta_array
  ! theta_array src_u src_v src_w adv_u adv_v adv_w.  n returns the length.
  function fel_drv_field(name, n) bind(C, name="fel_drv_field") result(ptr)
    character(kind=c_char), intent(in) :: name(*)
    integer(c_int64_t), intent(out) :: n
    type(c_ptr) :: ptr
    character(len=32) :: f
    integer :: m
    f = ' '
    do m = 1, 32
      if (name(m) == c_null_char) exit
      f(m:m) = name(m)
    end do
    n = int(size(u), c_int64_t)
    ptr = c_null_ptr
    select case (trim(f))
    case ('u'); ptr = c_loc(u)
    case ('v'); ptr = c_loc(v)
    case ('w'); ptr = c_loc(w)
    case ('uo'); ptr = c_loc(uo)
    case ('vo'); ptr = c_loc(vo)
    case ('wo'); ptr = c_loc(wo)
    case ('un'); ptr = c_loc(un)
    case ('vn'); ptr = c_loc(vn)
    case ('wn'); ptr = c_loc(wn)
    case ('p'); ptr = c_loc(p)
    case ('po'); ptr = c_loc(po)
    case ('pn'); ptr = c_loc(pn)
    case ('dp'); ptr = c_loc(dp)
    case ('div'); ptr = c_loc(div)
    case ('phat'); ptr = c_loc(phat)
    case ('rho'); ptr = c_loc(rho)
    case ('rmu'); ptr = c_loc(rmu)
    case ('rhon'); ptr = c_loc(rhon)
    case ('rmun'); ptr = c_loc(rmun)
    case ('phi0'); ptr = c_loc(phi(-2,-2,-2,0))
    case ('phi1'); ptr = c_loc(phi(-2,-2,-2,1))
    case ('phin0'); ptr = c_loc(phin(-2,-2,-2,0))
    case ('phin1'); ptr = c_loc(phin(-2,-2,-2,1))
    case ('theta_0_array'); ptr = c_loc(theta_0_array)
    case ('l1_array'); ptr = c_loc(l1_array)
    case ('l2_array'); ptr = c_loc(l2_array)
    case ('zeta_array'); ptr = c_loc(zeta_array)
    case ('theta_array'); ptr = c_loc(theta_array)
    case ('src_u'); ptr = c_loc(src_u)
    case ('src_v'); ptr = c_loc(src_v)
    case ('src_w'); ptr = c_loc(src_w)
    case ('adv_u'); ptr = c_loc(adv_u)
    case ('adv_v'); ptr = c_loc(adv_v)
    case ('adv_w'); ptr = c_loc(adv_w)
    case ('adv_uo'); ptr = c_loc(adv_uo)
    case ('adv_vo'); ptr = c_loc(adv_vo)
    case ('adv_wo'); ptr = c_loc(adv_wo)
    case ('prs_u'); ptr = c_loc(prs_u)
    case ('prs_v'); ptr = c_loc(prs_v)
    case ('prs_w'); ptr = c_loc(prs_w)
    case ('vis_u'); ptr = c_loc(vis_u)
    case ('vis_v'); ptr = c_loc(vis_v)
    case ('vis_w'); ptr = c_loc(vis_w)
    case ('fst_u'); ptr = c_loc(fst_u)
    case ('fst_v'); ptr = c_loc(fst_v)
    case ('fst_w'); ptr = c_loc(fst_w)
    case ('fst_un'); ptr = c_loc(fst_un)
    case ('fst_vn'); ptr = c_loc(fst_vn)
    case ('fst_wn'); ptr = c_loc(fst_wn)
    case ('sum_fst_u'); ptr = c_loc(sum_fst_u)
    case ('sum_fst_v'); ptr = c_loc(sum_fst_v)
    case ('sum_fst_w'); ptr = c_loc(sum_fst_w)
    case ('sum_fst_un'); ptr = c_loc(sum_fst_un)
    case ('sum_fst_vn'); ptr = c_loc(sum_fst_vn)
    case ('sum_fst_wn'); ptr = c_loc(sum_fst_wn)
    case ('flphix'); ptr = c_loc(flphix)
    case ('flphiy'); ptr = c_loc(flphiy)
    case ('flphiz'); ptr = c_loc(flphiz)
    case ('phix'); ptr = c_loc(phix)
    case ('phiy'); ptr = c_loc(phiy)
    case ('phiz'); ptr = c_loc(phiz)
    case ('rn_pp'); ptr = c_loc(rn_pp); n = 3 * n
    case ('rn_12'); ptr = c_loc(rn_12); n = 3 * n
    case ('rn_13'); ptr = c_loc(rn_13); n = 3 * n
    case ('rn_23'); ptr = c_loc(rn_23); n = 3 * n
    case ('rnn_pp'); ptr = c_loc(rnn_pp); n = 3 * n
    case ('rnn_12'); ptr = c_loc(rnn_12); n = 3 * n
    case ('rnn_13'); ptr = c_loc(rnn_13); n = 3 * n
    case ('rnn_23'); ptr = c_loc(rnn_23); n = 3 * n
    case ('rkap'); ptr = c_loc(rkap); n = 3 * n
    case ('rkapn'); ptr = c_loc(rkapn); n = 3 * n
    case default; n = 0
    end select
  end function fel_drv_field

  ! Makes every array device-resident (uploading the fields the caller
  ! initialised) and runs the start-up boundary fills of
  ! srcs/main.f90:303-366 and the pressure coefficients of main.f90:467-468.
  ! Before fel_drv_start: flag /= 0 registers the pressure solve's modal
  ! arrays in the flow class (in the library's FP32 mode the solve then runs
  ! in FP32 with FP64 refinement); 0 (default) keeps them FP64.
  subroutine fel_drv_set_modes_flow_class(flag) bind(C, name="fel_drv_set_modes_flow_class")
    integer(c_int), value :: flag
    modes_flow_class = flag /= 0
  end subroutine fel_drv_set_modes_flow_class

  subroutine fel_drv_start() bind(C, name="fel_drv_start")
    integer(c_int64_t) :: n3
    integer :: m
    n3 = int(size(u), c_int64_t)
    call reg(u, n3, 1); call reg(v, n3, 1); call reg(w, n3, 1)
    ! uo, vo, wo, po start at zero as in srcs/init.f:24-41 (the host copies
    ! are not initialised: allocate does not clear reused heap memory)
    call reg(uo, n3, 0); call reg(vo, n3, 0); call reg(wo, n3, 0)
    call reg(un, n3, 0); call reg(vn, n3, 0); call reg(wn, n3, 0)
    call reg(p, n3, 1); call reg(po, n3, 0); call reg(pn, n3, 0)
    call reg(phat, n3, 0); call reg64(dp, n3, 0); call reg64(div, n3, 0)
    call reg(rho, n3, 0); call reg(rmu, n3, 0); call reg(rhon, n3, 0); call reg(rmun, n3, 0)
    call reg(adv_u, n3, 0); call reg(adv_v, n3, 0); call reg(adv_w, n3, 0)
    call reg(adv_uo, n3, 0); call reg(adv_vo, n3, 0); call reg(adv_wo, n3, 0)
    call reg(prs_u, n3, 0); call reg(prs_v, n3, 0); call reg(prs_w, n3, 0)
    call reg(vis_u, n3, 0); call reg(vis_v, n3, 0); call reg(vis_w, n3, 0)
    call reg(fst_u, n3, 0); call reg(fst_v, n3, 0); call reg(fst_w, n3, 0)
    call reg(fst_un, n3, 0); call reg(fst_vn, n3, 0); call reg(fst_wn, n3, 0)
    call reg(sum_fst_u, n3, 0); call reg(sum_fst_v, n3, 0); call reg(sum_fst_w, n3, 0)
    call reg(sum_fst_un, n3, 0); call reg(sum_fst_vn, n3, 0); call reg(sum_fst_wn, n3, 0)
    call reg(src_u, n3, 0); call reg(src_v, n3, 0); call reg(src_w, n3, 0)
    call reg(phix, n3, 0); call reg(phiy, n3, 0); call reg(phiz, n3, 0)
    call reg64(flphix, n3, 0); call reg64(flphiy, n3, 0); call reg64(flphiz, n3, 0)
    ! each slice of phi and phin its own range, so that the step's slice
    ! copies (main.f90:695-697) are deferred like the other rotations
    do m = 0, nbub
      call reg64(phi(-2,-2,-2,m), n3, 1); call reg64(phin(-2,-2,-2,m), n3, 0)
    end do
    call reg(s, 6*n3, 0); call reg(tau, 6*n3, 0)
    call reg(rn_pp, 3*n3, 0); call reg(rn_12, 3*n3, 0); call reg(rn_13, 3*n3, 0); call reg(rn_23, 3*n3, 0)
    call reg(rnn_pp, 3*n3, 0); call reg(rnn_12, 3*n3, 0); call reg(rnn_13, 3*n3, 0); call reg(rnn_23, 3*n3, 0)
    call reg(rkap, 3*n3, 0); call reg(rkapn, 3*n3, 0)
    do m = 1, 45
      call reg(cf(-2,-2,-2,m), n3, 0)
    end do
    call reg64(theta_0_array, n3, 1); call reg64(l1_array, n3, 1); call reg64(l2_array, n3, 1)
    call reg64(zeta_array, n3, 1); call reg64(theta_array, n3, 0)
    if (modes_flow_class) then
      call reg(phir_r, int(ni,c_int64_t)*nj*nk, 0); call reg(phir_i, int(ni,c_int64_t)*nj*nk, 0)
    else
      call reg64(phir_r, int(ni,c_int64_t)*nj*nk, 0); call reg64(phir_i, int(ni,c_int64_t)*nj*nk, 0)
    end if

    ! srcs/main.f90:304-366
    call cpy(ni,nj,nk,u,un)
    call cpy(ni,nj,nk,v,vn)
    call cpy(ni,nj,nk,w,wn)
    call gnbc(nID, ni, nj, nk, u, w, c%uwall, theta_0_array, c%surface_tension, zeta_array, theta_array)
    call bnd_contact_angle(nID,ni,nj,nk,phi(-2,-2,-2,1),theta_array,c%dx,c%dy,c%dz)
    call bnd_periodic(ni,nj,nk,phi(-2,-2,-2,1))
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,phi(-2,-2,-2,1))
    call summation(ni,nj,nk,phi,nbub)
    call bndu(nID,ni,nj,nk,u ,v ,w ,c%uwall,c%dy,l1_array,l2_array,phi(-2,-2,-2,1))
    call bndu(nID,ni,nj,nk,uo,vo,wo,c%uwall,c%dy,l1_array,l2_array,phi(-2,-2,-2,1))
    call bndu(nID,ni,nj,nk,un,vn,wn,c%uwall,c%dy,l1_array,l2_array,phi(-2,-2,-2,1))
    call bnd_periodic(ni,nj,nk,u ); call bnd_periodic(ni,nj,nk,v ); call bnd_periodic(ni,nj,nk,w )
    call bnd_periodic(ni,nj,nk,uo); call bnd_periodic(ni,nj,nk,vo); call bnd_periodic(ni,nj,nk,wo)
    call bnd_periodic(ni,nj,nk,un); call bnd_periodic(ni,nj,nk,vn); call bnd_periodic(ni,nj,nk,wn)
    call gnbc(nID, ni, nj, nk, u, w, c%uwall, theta_0_array, c%surface_tension, zeta_array, theta_array)
    call bnd_contact_angle(nID,ni,nj,nk,phi(-2,-2,-2,0),theta_array,c%dx,c%dy,c%dz)
    call bnd_periodic(ni,nj,nk,phi(-2,-2,-2,0))
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,phi(-2,-2,-2,0))

    ! srcs/main.f90:467-468
    call cal_coef_prs(svall(2),c%rhog,dxinv,dyinv,dzinv,aw_p,ae_p,as_p,an_p,ab_p,at_p,ap_p)
    call cal_coef_solp_fft(svall(2),as_p,an_p,as_p_fft,an_p_fft,ap_p_fft)
    nstep = 0
    time = 0.0d0
  end subroutine fel_drv_start

  ! One timestep, srcs/main.f90:473-748 without the output cadences.
  subroutine step1()
    integer :: l
    real(8) :: rho_av, rhon_av, avphin
    nstep = nstep + 1
    if (c%cfl > 0.0d0) then
      ! srcs/main.f90:479 (commented out in the reference, which runs a fixed dt)
      call caldt(ipara,nID,ID,ndiv,ni,nj,nk,nstep,c%imon_t,dxinv,dyinv,dzinv,c%cfl,c%rhol,c%rhog, &
                 c%rmul,c%rmug,c%surface_tension,u,v,w,dt,time)
    else
      dt = c%dt
      time = time + dt
    end if
    l = 1
    ! VOF (main.f90:489-504)
    call cal_grad_p2a(ID,svall(2),ni,nj,nk,dxinv,dyinv_array,dzinv,phi(-2,-2,-2,l),phix,phiy,phiz)
    call solphi_mthinc1(ni,nj,nk,dxinv,dyinv,dzinv,dt,c%bet_mthinc,phi(-2,-2,-2,l),phix,phiy,phiz, &
                        u,v,w,flphix,flphiy,flphiz)
    call solphi_mthinc2(ni,nj,nk,dxinv,dyinv,dzinv,dt,u,v,w,flphix,flphiy,flphiz, &
                        phi(-2,-2,-2,l),phin(-2,-2,-2,l))
    call cal_grad_p2a(ID,svall(2),ni,nj,nk,dxinv,dyinv_array,dzinv,phin(-2,-2,-2,l),phix,phiy,phiz)
    call solphi_mthinc3(ipara,ni,nj,nk,dxinv,dyinv,dzinv,c%bet_mthinc,phix,phiy,phiz, &
                        phi(-2,-2,-2,l),phin(-2,-2,-2,l))
    call gnbc(nID, ni, nj, nk, u, w, c%uwall, theta_0_array, c%surface_tension, zeta_array, theta_array)
    call bnd_contact_angle(nID,ni,nj,nk,phin(-2,-2,-2,l),theta_array,c%dx,c%dy,c%dz)
    call bnd_periodic(ni,nj,nk,phin(-2,-2,-2,l))
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,phin(-2,-2,-2,l))
    ! main.f90:509-524; the contact-angle ghosts go to the old phi(0) (A3)
    call summation(ni,nj,nk,phin,nbub)
    call bnd_neumann(nID,ni,nj,nk,phin(-2,-2,-2,0))
    call gnbc(nID, ni, nj, nk, u, w, c%uwall, theta_0_array, c%surface_tension, zeta_array, theta_array)
    call bnd_contact_angle(nID,ni,nj,nk,phi(-2,-2,-2,0),theta_array,c%dx,c%dy,c%dz)
    call bnd_periodic(ni,nj,nk,phin(-2,-2,-2,0))
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,phin(-2,-2,-2,0))
    if (c%imon_t > 0) then
      if (mod(nstep,c%imon_t) == 0) then
        call cal_av(ipara,ndiv,ni,nj,nk,phin,avphin)
        write(*,'("AV_PHI   ",1i10,20e20.10)') nstep, avphin
      end if
    end if
    ! properties (main.f90:526-529)
    call cal_arith_mean(ni,nj,nk,phi ,c%rhol,c%rhog,rho )
    call cal_arith_mean(ni,nj,nk,phin,c%rhol,c%rhog,rhon)
    call cal_arith_mean(ni,nj,nk,phi ,c%rmul,c%rmug,rmu )
    call cal_arith_mean(ni,nj,nk,phin,c%rmul,c%rmug,rmun)
    ! explicit terms (main.f90:532-541)
    call cal_advu_weno5(ni,nj,nk,dxinv,dyinv,dzinv,u ,v ,w ,adv_u ,adv_v ,adv_w )
    call cal_advu_weno5(ni,nj,nk,dxinv,dyinv,dzinv,uo,vo,wo,adv_uo,adv_vo,adv_wo)
    call calphat(ni,nj,nk,p,po,phat)
    call cal_grad_prs(ni,nj,nk,dxinv,dyinv,dzinv,phat,prs_u,prs_v,prs_w)
    call calsij(ni,nj,nk,dxinv,dyinv,dzinv,u ,v ,w ,s )
    call cal_arith_tau(ni,nj,nk,s,rmu,tau)
    call cal_div_tensor(ni,nj,nk,dxinv,dyinv,dzinv,tau ,vis_u ,vis_v ,vis_w )
    ! CSF (main.f90:545-573)
    call init_q(ni,nj,nk,sum_fst_u,sum_fst_v,sum_fst_w)
    call init_q(ni,nj,nk,sum_fst_un,sum_fst_vn,sum_fst_wn)
    call cal_grad_p2a(ID,svall(2),ni,nj,nk,dxinv,dyinv_array,dzinv,phi(-2,-2,-2,l),phix,phiy,phiz)
    call calrn(ni,nj,nk,dxinv,dyinv,dzinv,1,1,1,phix,phiy,phiz,rn_pp)
    call calrn(ni,nj,nk,dxinv,dyinv,dzinv,0,0,1,phix,phiy,phiz,rn_12)
    call calrn(ni,nj,nk,dxinv,dyinv,dzinv,0,1,0,phix,phiy,phiz,rn_13)
    call calrn(ni,nj,nk,dxinv,dyinv,dzinv,1,0,0,phix,phiy,phiz,rn_23)
    call cal_grad_p2a(ID,svall(2),ni,nj,nk,dxinv,dyinv_array,dzinv,phin(-2,-2,-2,l),phix,phiy,phiz)
    call calrn(ni,nj,nk,dxinv,dyinv,dzinv,1,1,1,phix,phiy,phiz,rnn_pp)
    call calrn(ni,nj,nk,dxinv,dyinv,dzinv,0,0,1,phix,phiy,phiz,rnn_12)
    call calrn(ni,nj,nk,dxinv,dyinv,dzinv,0,1,0,phix,phiy,phiz,rnn_13)
    call calrn(ni,nj,nk,dxinv,dyinv,dzinv,1,0,0,phix,phiy,phiz,rnn_23)
    call calrkap(ni,nj,nk,dxinv,dyinv,dzinv, rn_pp, rn_12, rn_13, rn_23,rkap )
    call calrkap(ni,nj,nk,dxinv,dyinv,dzinv,rnn_pp,rnn_12,rnn_13,rnn_23,rkapn)
    call cal_grad_p2uvw(ID,svall(2),ni,nj,nk,dxinv,dyinv_array,dzinv,phi(-2,-2,-2,l),phix,phiy,phiz)
    call calfst(ni,nj,nk,c%rhol,c%rhog,c%surface_tension,phix,phiy,phiz,rkap,fst_u,fst_v,fst_w)
    call cal_grad_p2uvw(ID,svall(2),ni,nj,nk,dxinv,dyinv_array,dzinv,phin(-2,-2,-2,l),phix,phiy,phiz)
    call calfst(ni,nj,nk,c%rhol,c%rhog,c%surface_tension,phix,phiy,phiz,rkapn,fst_un,fst_vn,fst_wn)
    call sum_fst(ni,nj,nk,fst_u,fst_v,fst_w,sum_fst_u,sum_fst_v,sum_fst_w)
    call sum_fst(ni,nj,nk,fst_un,fst_vn,fst_wn,sum_fst_un,sum_fst_vn,sum_fst_wn)
    call cal_av(ipara,ndiv,ni,nj,nk,rho ,rho_av)
    call cal_av(ipara,ndiv,ni,nj,nk,rhon,rhon_av)
    rho_av = (rho_av+rhon_av)*0.5d0
    ! source (main.f90:579-588)
    call cal_srcu(nID,ni,nj,nk,rho_av,dt,c%grvb,c%grvp,u,v,w,rho,rhon, &
                  adv_u,adv_v,adv_w,adv_uo,adv_vo,adv_wo,prs_u,prs_v,prs_w,vis_u,vis_v,vis_w, &
                  sum_fst_u,sum_fst_v,sum_fst_w,sum_fst_un,sum_fst_vn,sum_fst_wn,src_u,src_v,src_w)
    ! semi-implicit viscous step (main.f90:590-631)
    call cal_arith_coef_vis(ni,nj,nk,dxinv,dyinv,dzinv,rmun, &
      cf(-2,-2,-2, 1),cf(-2,-2,-2, 2),cf(-2,-2,-2, 3),cf(-2,-2,-2, 4),cf(-2,-2,-2, 5), &
      cf(-2,-2,-2, 6),cf(-2,-2,-2, 7),cf(-2,-2,-2, 8),cf(-2,-2,-2, 9),cf(-2,-2,-2,10), &
      cf(-2,-2,-2,11),cf(-2,-2,-2,12),cf(-2,-2,-2,13),cf(-2,-2,-2,14),cf(-2,-2,-2,15), &
      cf(-2,-2,-2,16),cf(-2,-2,-2,17),cf(-2,-2,-2,18),cf(-2,-2,-2,19),cf(-2,-2,-2,20), &
      cf(-2,-2,-2,21),cf(-2,-2,-2,22),cf(-2,-2,-2,23),cf(-2,-2,-2,24),cf(-2,-2,-2,25), &
      cf(-2,-2,-2,26),cf(-2,-2,-2,27),cf(-2,-2,-2,28),cf(-2,-2,-2,29),cf(-2,-2,-2,30), &
      cf(-2,-2,-2,31),cf(-2,-2,-2,32),cf(-2,-2,-2,33),cf(-2,-2,-2,34),cf(-2,-2,-2,35), &
      cf(-2,-2,-2,36),cf(-2,-2,-2,37),cf(-2,-2,-2,38),cf(-2,-2,-2,39),cf(-2,-2,-2,40), &
      cf(-2,-2,-2,41),cf(-2,-2,-2,42),cf(-2,-2,-2,43),cf(-2,-2,-2,44),cf(-2,-2,-2,45))
    call cpy(ni,nj,nk,u,un)
    call cpy(ni,nj,nk,v,vn)
    call cpy(ni,nj,nk,w,wn)
    call solu_sor4(ipara,ID,nID,ndiv,ni,nj,nk,key,sendjb,recvjb,nstep,c%imon_t,dt,rho,rhon, &
      cf(-2,-2,-2, 1),cf(-2,-2,-2, 2),cf(-2,-2,-2, 3),cf(-2,-2,-2, 4),cf(-2,-2,-2, 5), &
      cf(-2,-2,-2, 6),cf(-2,-2,-2, 7),cf(-2,-2,-2, 8),cf(-2,-2,-2, 9),cf(-2,-2,-2,10), &
      cf(-2,-2,-2,11),cf(-2,-2,-2,12),cf(-2,-2,-2,13),cf(-2,-2,-2,14),cf(-2,-2,-2,15), &
      cf(-2,-2,-2,16),cf(-2,-2,-2,17),cf(-2,-2,-2,18),cf(-2,-2,-2,19),cf(-2,-2,-2,20), &
      cf(-2,-2,-2,21),cf(-2,-2,-2,22),cf(-2,-2,-2,23),cf(-2,-2,-2,24),cf(-2,-2,-2,25), &
      cf(-2,-2,-2,26),cf(-2,-2,-2,27),cf(-2,-2,-2,28),cf(-2,-2,-2,29),cf(-2,-2,-2,30), &
      cf(-2,-2,-2,31),cf(-2,-2,-2,32),cf(-2,-2,-2,33),cf(-2,-2,-2,34),cf(-2,-2,-2,35), &
      cf(-2,-2,-2,36),cf(-2,-2,-2,37),cf(-2,-2,-2,38),cf(-2,-2,-2,39),cf(-2,-2,-2,40), &
      cf(-2,-2,-2,41),cf(-2,-2,-2,42),cf(-2,-2,-2,43),cf(-2,-2,-2,44),cf(-2,-2,-2,45), &
      src_u,src_v,src_w,un,vn,wn,c%uwall,c%dy,l1_array,l2_array,phi(-2,-2,-2,1))
    call bndu(nID,ni,nj,nk,un,vn,wn,c%uwall,c%dy,l1_array,l2_array,phi(-2,-2,-2,1))
    call bnd_periodic(ni,nj,nk,un)
    call bnd_periodic(ni,nj,nk,vn)
    call bnd_periodic(ni,nj,nk,wn)
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,un)
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,vn)
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,wn)
    ! projection (main.f90:633-661)
    call caldiv(ipara,ndiv,ni,nj,nk,dxinv,dyinv,dzinv,dt,un,vn,wn,div,err_div0)
    call solp_fft_tdma1(ni,nj,nk,fftdata,div,phir_r,phir_i)
    call trans_r2w(ipara,ID,ndiv,svall,key,sendbuf,recvbuf,phir_r,phir_r)
    call trans_r2w(ipara,ID,ndiv,svall,key,sendbuf,recvbuf,phir_i,phir_i)
    call solp_fft_tdma2(ID,ndiv,svall,c%rhog,dxinv,dzinv,as_p_fft,an_p_fft,ap_p_fft,atdma, &
                        btdma_r,btdma_i,phir_r,phir_i)
    call trans_w2r(ipara,ID,ndiv,svall,key,sendbuf,recvbuf,phir_r,phir_r)
    call trans_w2r(ipara,ID,ndiv,svall,key,sendbuf,recvbuf,phir_i,phir_i)
    call solp_fft_tdma3(ni,nj,nk,fftdata,phir_r,phir_i,dp)
    call bnd_avzero(ipara,ndiv,ni,nj,nk,dp)
    call bnd_neumann(nID,ni,nj,nk,dp)
    call bnd_periodic(ni,nj,nk,dp)
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,dp)
    call solp_fft_tdma4(ipara,ID,ndiv,ni,nj,nk,nstep,c%imon_t,c%rhog,dxinv,dyinv,dzinv,div,dp)
    call corunp_explicit(nID,ni,nj,nk,c%rhog,dxinv,dyinv,dzinv,dt,dp,phat,un,vn,wn,pn)
    call bndu(nID,ni,nj,nk,un,vn,wn,c%uwall,c%dy,l1_array,l2_array,phi(-2,-2,-2,1))
    call bnd_periodic(ni,nj,nk,un)
    call bnd_periodic(ni,nj,nk,vn)
    call bnd_periodic(ni,nj,nk,wn)
    call bnd_avzero(ipara,ndiv,ni,nj,nk,pn)
    call bnd_neumann(nID,ni,nj,nk,pn)
    call bnd_periodic(ni,nj,nk,pn)
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,pn)
    ! diagnostics (main.f90:663-684)
    call cal_center(ipara,ID,svall(2),ni,nj,nk,c%xl,c%yl,c%zl,dt,c%rhog,phi(-2,-2,-2,1), &
                    center,center_pre1,center_pre2,area_pre1,area_pre2,lap,velocity)
    call cal_vel(ipara,ni,nj,nk,c%particle_radius,c%rhol,c%rmul,u,v,w,phi(-2,-2,-2,1),vel,re)
    if (c%imon_t > 0) then
      if (mod(nstep,c%imon_t) == 0) then
        call caldiv(ipara,ndiv,ni,nj,nk,dxinv,dyinv,dzinv,dt,un,vn,wn,div,err_div)
        write(*,'("Err_div  ",1i10,20e20.10)') nstep, err_div/err_div0, err_div, err_div0
      end if
    end if
    ! state rotation (main.f90:686-697)
    call cpy(ni,nj,nk,u ,uo)
    call cpy(ni,nj,nk,v ,vo)
    call cpy(ni,nj,nk,w ,wo)
    call cpy(ni,nj,nk,p ,po)
    call cpy(ni,nj,nk,un,u )
    call cpy(ni,nj,nk,vn,v )
    call cpy(ni,nj,nk,wn,w )
    call cpy(ni,nj,nk,pn,p )
    call cpy(ni,nj,nk,phin(-2,-2,-2,0),phi(-2,-2,-2,0))
    call cpy(ni,nj,nk,phin(-2,-2,-2,1),phi(-2,-2,-2,1))
  end subroutine step1

  subroutine fel_drv_steps(n) bind(C, name="fel_drv_steps")
    integer(c_int), value :: n
    integer :: m
    do m = 1, n
      call step1()
    end do
    call flush(6)
  end subroutine fel_drv_steps

  ! Device -> host copy of a field listed in fel_drv_field.
  subroutine fel_drv_download(name) bind(C, name="fel_drv_download")
    character(kind=c_char), intent(in) :: name(*)
    integer(c_int64_t) :: n
    type(c_ptr) :: ptr
    ptr = fel_drv_field(name, n)
    if (c_associated(ptr)) call felgpu_download_ptr(ptr)
  end subroutine fel_drv_download

  subroutine felgpu_download_ptr(ptr)
    type(c_ptr), value :: ptr
    interface
      subroutine felgpu_download_c(h) bind(C, name="felgpu_download_")
        import :: c_ptr
        type(c_ptr), value :: h
      end subroutine felgpu_download_c
    end interface
    call felgpu_download_c(ptr)
  end subroutine felgpu_download_ptr

  ! ---------------------------------------------------------------------
  ! Checkpoint/restart in the reference's file format.  srcs/dataou.f:24-38
  ! and srcs/datain.f:24-38 write/read ONE unformatted sequential record:
  ! the padded images of u, v, w, p, uo, vo, wo, po (column-major, ghosts
  ! included), phi(...,0:nbub), time (real*8) and nstep (default integer).
  ! A whole-array I/O list transfers the same elements in the same order as
  ! the reference's implied-do loops, so the files are interchangeable with
  ! the reference's (checked against its own dataou/datain).
  ! ---------------------------------------------------------------------

  ! Writes one checkpoint record from host arrays (no device involved).
  subroutine fel_ckpt_write(path, plen, ci, cj, ck, cb, step, t, a1, a2, a3, a4, a5, a6, a7, a8, &
                            ph) bind(C, name="fel_ckpt_write")
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int), value :: plen, ci, cj, ck, cb
    integer(c_int), intent(in) :: step
    real(c_double), intent(in) :: t
    real(c_double), intent(in), dimension(-2:ci+3,-2:cj+3,-2:ck+3) :: a1, a2, a3, a4, a5, a6, a7, a8
    real(c_double), intent(in) :: ph(-2:ci+3,-2:cj+3,-2:ck+3,0:cb)
    integer :: lun, ios
    open(newunit=lun, file=c_path(path, plen), form='unformatted', access='sequential', &
         status='replace', action='write', iostat=ios)
    if (ios /= 0) error stop 'fel_ckpt_write: cannot open the checkpoint file'
    write(lun) a1, a2, a3, a4, a5, a6, a7, a8, ph, t, step
    close(lun)
  end subroutine fel_ckpt_write

  ! Reads one checkpoint record into host arrays (no device involved).
  subroutine fel_ckpt_read(path, plen, ci, cj, ck, cb, step, t, a1, a2, a3, a4, a5, a6, a7, a8, &
                           ph) bind(C, name="fel_ckpt_read")
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int), value :: plen, ci, cj, ck, cb
    integer(c_int), intent(out) :: step
    real(c_double), intent(out) :: t
    real(c_double), intent(out), dimension(-2:ci+3,-2:cj+3,-2:ck+3) :: a1, a2, a3, a4, a5, a6, a7, a8
    real(c_double), intent(out) :: ph(-2:ci+3,-2:cj+3,-2:ck+3,0:cb)
    integer :: lun, ios
    open(newunit=lun, file=c_path(path, plen), form='unformatted', access='sequential', &
         status='old', action='read', iostat=ios)
    if (ios /= 0) error stop 'fel_ckpt_read: cannot open the checkpoint file'
    read(lun, iostat=ios) a1, a2, a3, a4, a5, a6, a7, a8, ph, t, step
    if (ios /= 0) error stop 'fel_ckpt_read: the record does not match this grid'
    close(lun)
  end subroutine fel_ckpt_read

  function c_path(path, plen) result(f)
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int), intent(in) :: plen
    character(len=plen) :: f
    integer :: m
    do m = 1, plen
      f(m:m) = path(m)
    end do
  end function c_path

  ! srcs/main.f90:702-704: the run's state to a checkpoint file (device ->
  ! host copies of the nine arrays first).
  subroutine fel_drv_dataou(path, plen) bind(C, name="fel_drv_dataou")
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int), value :: plen
    integer :: l
    call felgpu_download_ptr(c_loc(u)); call felgpu_download_ptr(c_loc(v))
    call felgpu_download_ptr(c_loc(w)); call felgpu_download_ptr(c_loc(p))
    call felgpu_download_ptr(c_loc(uo)); call felgpu_download_ptr(c_loc(vo))
    call felgpu_download_ptr(c_loc(wo)); call felgpu_download_ptr(c_loc(po))
    do l = 0, nbub
      call felgpu_download_ptr(c_loc(phi(-2,-2,-2,l)))
    end do
    call fel_ckpt_write(path, plen, ni, nj, nk, nbub, nstep, time, u, v, w, p, uo, vo, wo, po, phi)
  end subroutine fel_drv_dataou

  ! srcs/main.f90:375-404: restart of a started run from a checkpoint file:
  ! the nine arrays, time and nstep are replaced (and uploaded), then the
  ! reference's restart fills refresh the ghosts -- bndu of (u,v,w) and
  ! (uo,vo,wo), bnd_periodic of the eight fields, gnbc and the contact-angle
  ! ghosts of phi(...,0).
  subroutine fel_drv_datain(path, plen) bind(C, name="fel_drv_datain")
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int), value :: plen
    integer :: l
    call fel_ckpt_read(path, plen, ni, nj, nk, nbub, nstep, time, u, v, w, p, uo, vo, wo, po, phi)
    call felgpu_upload(u); call felgpu_upload(v); call felgpu_upload(w)
    call felgpu_upload(p); call felgpu_upload(uo); call felgpu_upload(vo)
    call felgpu_upload(wo); call felgpu_upload(po)
    do l = 0, nbub
      call felgpu_upload(phi(-2,-2,-2,l))
    end do
    call bndu(nID,ni,nj,nk,u ,v ,w ,c%uwall,c%dy,l1_array,l2_array,phi(-2,-2,-2,1))
    call bndu(nID,ni,nj,nk,uo,vo,wo,c%uwall,c%dy,l1_array,l2_array,phi(-2,-2,-2,1))
    call bnd_periodic(ni,nj,nk,u ); call bnd_periodic(ni,nj,nk,v ); call bnd_periodic(ni,nj,nk,w )
    call bnd_periodic(ni,nj,nk,p ); call bnd_periodic(ni,nj,nk,uo); call bnd_periodic(ni,nj,nk,vo)
    call bnd_periodic(ni,nj,nk,wo); call bnd_periodic(ni,nj,nk,po)
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,u )
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,v )
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,w )
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,p )
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,uo)
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,vo)
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,wo)
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,po)
    call gnbc(nID, ni, nj, nk, u, w, c%uwall, theta_0_array, c%surface_tension, zeta_array, theta_array)
    call bnd_contact_angle(nID,ni,nj,nk,phi(-2,-2,-2,0),theta_array,c%dx,c%dy,c%dz)
    call bnd_periodic(ni,nj,nk,phi(-2,-2,-2,0))
    call bnd_comm(ipara,nID,ni,nj,nk,key,sendjb,recvjb,phi(-2,-2,-2,0))
  end subroutine fel_drv_datain

  ! ---------------------------------------------------------------------
  ! Output cadences of srcs/main.f90:706-743 and 663-672.  The device
  ! reduces the fields (felgpu_vtk_points_, mk_slice_j_,
  ! felgpu_interface_crossings_); these routines format the reference's
  ! files and monitor lines with the same Fortran edit descriptors.
  ! ---------------------------------------------------------------------

  ! The VTK file of srcs/mkvtk_phi.f:55-83 (and mkvtk_p.f) from the vertex
  ! values vals((mi+1)(mj+1)(mk+1), i fastest): header, then one value per
  ! line in E10.3.
  subroutine fel_vtk_write(path, plen, mi, mj, mk, step, ddx, ddy, ddz, sname, slen, vals) &
      bind(C, name="fel_vtk_write")
    character(kind=c_char), intent(in) :: path(*), sname(*)
    integer(c_int), value :: plen, mi, mj, mk, step, slen
    real(c_double), value :: ddx, ddy, ddz
    real(c_double), intent(in) :: vals(*)
    integer :: lun, m, n
    n = (mi+1)*(mj+1)*(mk+1)
    open(newunit=lun, file=c_path(path, plen), status='replace')
    write(lun, "('# vtk DataFile Version 2.0')")
    write(lun, "('t=',I0)") step
    write(lun, "('ASCII')")
    write(lun, "('DATASET STRUCTURED_POINTS')")
    write(lun, "('DIMENSIONS ',I0,' ',I0,' ',I0)") mi+1, mj+1, mk+1
    write(lun, "('ORIGIN ',e11.4,' ',e11.4,' ',e11.4)") 0.0, 0.0, 0.0
    write(lun, "('ASPECT_RATIO ',e11.4,' ',e11.4,' ',e11.4)") ddx, ddy, ddz
    write(lun, "('')")
    write(lun, "('POINT_DATA ',I0)") n
    write(lun, "('SCALARS ',A,' float')") c_path(sname, slen)
    write(lun, "('LOOKUP_TABLE default')")
    do m = 1, n
      write(lun, "(e10.3)") vals(m)
    end do
    close(lun)
  end subroutine fel_vtk_write

  ! The monitor lines of srcs/find_interface_positions_upper.f90:21 (every
  ! crossing x with its z, kind 0) and _lower.f90:24-33 (the mean over k of
  ! the first crossing, 0 where a row has none, printed negated; kind 1),
  ! from the crossings x(ni, nk) / cnt(nk) of felgpu_interface_crossings_.
  subroutine fel_interface_print(kind, ci, ck, ddz, x, cnt) bind(C, name="fel_interface_print")
    integer(c_int), value :: kind, ci, ck
    real(c_double), value :: ddz
    real(c_double), intent(in) :: x(ci, ck)
    integer(c_int), intent(in) :: cnt(ck)
    integer :: k, m
    real(8) :: z, sum_x, x_k
    if (kind == 0) then
      do k = 1, ck
        z = (k - 0.5d0) * ddz
        do m = 1, cnt(k)
          print *, x(m, k), z
        end do
      end do
    else
      sum_x = 0.0d0
      do k = 1, ck
        x_k = 0.0d0
        if (cnt(k) > 0) x_k = x(1, k)
        sum_x = sum_x + x_k
      end do
      print *, 'Lower Interface Position (Average over k):'
      print *, -(sum_x / ck)
    end if
  end subroutine fel_interface_print

  ! Flushes the Fortran runtime's standard output (monitor lines).
  subroutine fel_drv_flush() bind(C, name="fel_drv_flush")
    flush(6)
  end subroutine fel_drv_flush

  ! srcs/main.f90:716-743 without the vorticity fields: phi_<nstep>.vtk and
  ! p_<nstep>.vtk in the current directory, then the interface positions on
  ! the top and bottom wall rows.  phi(...,0) is the field the reference
  ! gathers (phi_all) for these.
  subroutine fel_drv_vtk() bind(C, name="fel_drv_vtk")
    real(8), allocatable :: vals(:), xc(:)
    integer(c_int), allocatable :: cnt(:)
    character(len=32) :: fname
    integer(c_int) :: kind, jrow, first
    allocate(vals((ni+1)*(nj+1)*(nk+1)), xc(ni*nk), cnt(nk))
    kind = 0
    call felgpu_vtk_points(ni, nj, nk, phi, kind, vals)
    write(fname, '("phi_",i7.7,".vtk")') nstep
    call fel_vtk_write(fname, len_trim(fname), ni, nj, nk, nstep, c%dx, c%dy, c%dz, 'phi', 3, vals)
    kind = 1
    call felgpu_vtk_points(ni, nj, nk, p, kind, vals)
    write(fname, '("p_",i7.7,".vtk")') nstep
    call fel_vtk_write(fname, len_trim(fname), ni, nj, nk, nstep, c%dx, c%dy, c%dz, 'p', 1, vals)
    jrow = nj; first = 0
    call felgpu_interface_crossings(ni, nj, nk, phi, jrow, c%dx, c%xl, first, xc, cnt)
    call fel_interface_print(0, ni, nk, c%dz, xc, cnt)
    jrow = 1; first = 1
    call felgpu_interface_crossings(ni, nj, nk, phi, jrow, c%dx, c%xl, first, xc, cnt)
    call fel_interface_print(1, ni, nk, c%dz, xc, cnt)
    deallocate(vals, xc, cnt)
  end subroutine fel_drv_vtk

  ! srcs/main.f90:706-713 with srcs/mkuwp.h: the slices of u, w and phi at
  ! j = particle_radius/dy and the file uwphi<nstep> (x, z, u, w, phi per
  ! line in 20E20.10, values below 1e-40 in magnitude written as 0, a blank
  ! line after each k).
  subroutine fel_drv_uwphi() bind(C, name="fel_drv_uwphi")
    real(8), allocatable :: us(:,:), ws(:,:), ps(:,:), wrk(:,:)
    character(len=32) :: fname
    integer :: j, i, k, lun
    real(8) :: u00, w00, q00
    allocate(us(0:ni,0:nk), ws(0:ni,0:nk), ps(0:ni,0:nk), wrk(0:ni,0:nk))
    us = 0.0d0; ws = 0.0d0; ps = 0.0d0
    j = int(c%particle_radius / c%dy)
    call mk_slice_j(ipara,ID,nID,ndiv,ni,nj,nk,j,0,1,1,wrk,u,us)
    call mk_slice_j(ipara,ID,nID,ndiv,ni,nj,nk,j,1,1,0,wrk,w,ws)
    call mk_slice_j(ipara,ID,nID,ndiv,ni,nj,nk,j,1,1,1,wrk,phi,ps)
    write(fname, '("uwphi",i7.7)') nstep
    open(newunit=lun, file=trim(fname))
    do k = 0, nk
      do i = 0, ni
        u00 = us(i,k); w00 = ws(i,k); q00 = ps(i,k)
        if (abs(u00) < 1.0d-40) u00 = 0.0d0
        if (abs(w00) < 1.0d-40) w00 = 0.0d0
        if (abs(q00) < 1.0d-40) q00 = 0.0d0
        write(lun, '(20e20.10)') dble(i)*c%dx, dble(k)*c%dz, u00, w00, q00
      end do
      write(lun, '()')
    end do
    close(lun)
    deallocate(us, ws, ps, wrk)
  end subroutine fel_drv_uwphi

  ! srcs/mk_center.f and srcs/mk_vel.f for the last step's cal_center and
  ! cal_vel results: one line appended to center_01.dat and velocity_01.dat.
  subroutine fel_drv_center_vel() bind(C, name="fel_drv_center_vel")
    integer :: lun
    open(newunit=lun, file='center_01.dat', status='unknown', position='append')
    write(lun, '(20e20.10)') time, center(1), center(2), center(3), velocity, &
      sqrt((center(1)-0.5d0)**2 + (center(3)-0.5d0)**2)
    close(lun)
    open(newunit=lun, file='velocity_01.dat', status='unknown', position='append')
    write(lun, '(20e20.10)') time, vel(1), vel(2), vel(3), re(1), re(2), velocity
    close(lun)
  end subroutine fel_drv_center_vel

  ! Step counter, time, last SOR iteration count and residual ratio, the
  ! pre-projection divergence RMS, the diagnostics of cal_center/cal_vel and
  ! the last step's dt.
  subroutine fel_drv_state(out) bind(C, name="fel_drv_state")
    real(c_double), intent(out) :: out(13)
    integer(c_int) :: lp
    real(c_double) :: ratio
    call felgpu_sor_last(lp, ratio)
    out(1) = nstep; out(2) = time; out(3) = lp; out(4) = ratio; out(5) = err_div0
    out(6:8) = center; out(9) = velocity; out(10:12) = vel
    out(13) = dt
  end subroutine fel_drv_state

  subroutine fel_drv_destroy() bind(C, name="fel_drv_destroy")
    if (.not. created) return
    call felgpu_unregister_all()
    deallocate(u, v, w, uo, vo, wo, un, vn, wn, p, po, pn, phat, dp, div, rho, rmu, rhon, rmun)
    deallocate(adv_u, adv_v, adv_w, adv_uo, adv_vo, adv_wo, prs_u, prs_v, prs_w, vis_u, vis_v, vis_w)
    deallocate(fst_u, fst_v, fst_w, fst_un, fst_vn, fst_wn, sum_fst_u, sum_fst_v, sum_fst_w)
    deallocate(sum_fst_un, sum_fst_vn, sum_fst_wn, src_u, src_v, src_w, phix, phiy, phiz)
    deallocate(flphix, flphiy, flphiz, phi, phin, s, tau, rn_pp, rn_12, rn_13, rn_23)
    deallocate(rnn_pp, rnn_12, rnn_13, rnn_23, rkap, rkapn, cf)
    deallocate(theta_0_array, l1_array, l2_array, zeta_array, theta_array, phir_r, phir_i)
    deallocate(dyinv_array, fftdata, aw_p, ae_p, as_p, an_p, ab_p, at_p, ap_p)
    deallocate(as_p_fft, an_p_fft, ap_p_fft, atdma, btdma_r, btdma_i)
    created = .false.
  end subroutine fel_drv_destroy

end module fel_driver
