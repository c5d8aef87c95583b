! Glue between the reference program's main loop and libfsil.so, for the
! drop-in check (tests/test_gpu_fortran_host.py).  build_host.py compiles the
! reference's own sources with four call sites of main.f90 redirected here:
!
!   main.f90:62  start-up  -> fsil_host_init   (blocks and bodies onto the device)
!   main.f90:106 the step  -> fsil_host_step   (replaces tree_collision_streaming_IBM_FEM)
!   main.f90:107 macros    -> fsil_host_macro  (den/uuu back for the writers)
!   main.f90:118 continue  -> fsil_host_fin    (fIn back before write_continue_blocks)
!
! The fluid step (volume force, macros, collision, streaming, boundary faces)
! and the immersed-boundary coupling (UpdateElmtInterp_, the multi-direct
! forcing loop, spreading and the beam loads) run on the device through
! include/fsil_bridge.f90; the bodies' markers (UpdatePosVelArea), their
! structure solver (Solver) and every output writer are the host program's
! own, fed with the device's marker forces and beam loads.  Single blocks
! (block trees stop with a message).
module fsil_host
    use iso_c_binding
    use fsil_bridge
    use FlowCondition, only: flow
    use SolidBody, only: VBodies, Solver
    use FluidDomain, only: LBMblks
    use LBMBlockComm, only: blockTree
    implicit none
    private
    public :: fsil_host_init, fsil_host_step, fsil_host_macro, fsil_host_fin
    integer(c_int), allocatable :: hnd(:)       ! device block handle of each block
    real(c_double), allocatable :: Exyz(:, :), Evel(:, :), Ea(:), xc(:, :), Eforce(:, :), lod(:)
contains

    ! sizes of the bodies carried by block ib: markers, nodes and elements
    subroutine carried_sizes(ib, nmark, nelem, nnode)
        integer, intent(in) :: ib
        integer, intent(out) :: nmark, nelem, nnode
        integer :: i, b
        nmark = 0; nelem = 0; nnode = 0
        do i = 1, LBMblks(ib)%carriedBodies(0)
            b = LBMblks(ib)%carriedBodies(i)
            nmark = nmark + VBodies(b)%v_nelmts
            nelem = nelem + VBodies(b)%rbm%nEL
            nnode = nnode + VBodies(b)%rbm%nND
        enddo
    end subroutine

    ! after check_is_continue and before the start-up boundary update
    ! (main.f90:62-64): every block onto the device with the state the host
    ! has read or initialised, its volume force and boundary conditions set
    ! there, and the bodies it carries registered for the coupling
    subroutine fsil_host_init()
        integer :: ib, i, b, e, m, nb, nmark, nelem, nnode, km, ke
        integer(c_int), allocatable :: nm(:), nn(:), ne(:), melem(:), enodes(:)
        integer(c_int) :: h
        call fsil_check(fsil_init(0), 'fsil_init')
        allocate(hnd(size(LBMblks)))
        do ib = 1, size(LBMblks)
            if (blockTree(ib)%nsons > 0) stop 'fsil_host: block trees are not wired in this host'
            associate (blk => LBMblks(ib))
                call fsil_check(fsil_block_create(int(blk%ID, c_int), [blk%xDim, blk%yDim, blk%zDim], &
                    [blk%dh, blk%xmin, blk%ymin, blk%zmin, blk%xmax, blk%ymax, blk%zmax], blk%BndConds, &
                    int(blk%iCollidModel, c_int), [blk%tau, blk%Omega, blk%Omega2], blk%M_COLLID, blk%M_FORCE, &
                    0_c_int, h), 'fsil_block_create')
                hnd(ib) = h
                call fsil_check(fsil_set_flow(h, flow%denIn, flow%uvwIn, flow%shearRateIn, &
                    int(flow%velocityKind, c_int), flow%volumeForceIn, flow%volumeForceAmp, flow%volumeForceFreq, &
                    flow%volumeForcePhi), 'fsil_set_flow')
                if (blk%iCollidModel == 11 .or. blk%iCollidModel == 14 .or. blk%iCollidModel == 15) &
                    call fsil_check(fsil_set_viscosity(h, flow%nu), 'fsil_set_viscosity')
                call fsil_check(fsil_f_upload(h, blk%fIn), 'fsil_f_upload')
                call fsil_check(fsil_set_time(h, blk%blktime), 'fsil_set_time')
                call fsil_check(fsil_set_boundary_conditions(h), 'fsil_set_boundary_conditions')
                nb = blk%carriedBodies(0)
                if (nb > 0) then
                    call carried_sizes(ib, nmark, nelem, nnode)
                    allocate(nm(nb), nn(nb), ne(nb), melem(nmark), enodes(2 * nelem))
                    km = 0; ke = 0
                    do i = 1, nb
                        b = blk%carriedBodies(i)
                        nm(i) = VBodies(b)%v_nelmts
                        nn(i) = VBodies(b)%rbm%nND
                        ne(i) = VBodies(b)%rbm%nEL
                        do m = 1, VBodies(b)%v_nelmts
                            melem(km + m) = VBodies(b)%vtor_f(m) - 1
                        enddo
                        do e = 1, VBodies(b)%rbm%nEL
                            enodes(ke + 2 * e - 1) = VBodies(b)%rbm%m_elements(e)%node0 - 1
                            enodes(ke + 2 * e) = VBodies(b)%rbm%m_elements(e)%node1 - 1
                        enddo
                        km = km + VBodies(b)%v_nelmts
                        ke = ke + 2 * VBodies(b)%rbm%nEL
                    enddo
                    call fsil_check(fsil_ibm_define_bodies(h, int(nb, c_int), nm, nn, ne, melem, enodes), &
                        'fsil_ibm_define_bodies')
                    call fsil_check(fsil_ibm_set_params(h, int(flow%ntolLBM, c_int), flow%dtolLBM, flow%Uref, &
                        flow%denIn), 'fsil_ibm_set_params')
                    deallocate(nm, nn, ne, melem, enodes)
                endif
            end associate
        enddo
    end subroutine

    ! tree_collision_streaming_IBM_FEM (LBMBlockComm.f90:279-318) of a single
    ! block: volume force, IBM_FEM (:320-338) with the coupling on the device
    ! and the host's markers and structure sub-steps, then collision,
    ! streaming and boundary faces on the device
    subroutine fsil_host_step(treenode, time_collision, time_streaming, time_IBM, time_FEM)
        integer, intent(in) :: treenode
        real(8), intent(inout) :: time_collision, time_streaming, time_IBM, time_FEM
        integer :: i, b, e, m, nb, nmark, nelem, nnode, km, ke, kn, isubstep
        integer(c_int) :: h, iters
        real(c_double) :: dmax
        real(8) :: dt_solid
        h = hnd(treenode)
        ! update_volume_force (LBMBlockComm.f90:283) on the device, and on the
        ! host for the writers that read the block's volume force
        call LBMblks(treenode)%update_volume_force()
        call fsil_check(fsil_set_time(h, LBMblks(treenode)%blktime), 'fsil_set_time')
        nb = LBMblks(treenode)%carriedBodies(0)
        if (nb > 0) then
            call carried_sizes(treenode, nmark, nelem, nnode)
            if (.not. allocated(Exyz)) allocate(Exyz(3, nmark), Evel(3, nmark), Ea(nmark), xc(3, nelem), &
                                                Eforce(3, nmark), lod(6 * nnode))
            do i = 1, nb
                call VBodies(LBMblks(treenode)%carriedBodies(i))%UpdatePosVelArea()
            enddo
            km = 0; ke = 0
            do i = 1, nb
                b = LBMblks(treenode)%carriedBodies(i)
                do m = 1, VBodies(b)%v_nelmts
                    Exyz(:, km + m) = VBodies(b)%v_Exyz(1:3, m)
                    Evel(:, km + m) = VBodies(b)%v_Evel(1:3, m)
                    Ea(km + m) = VBodies(b)%v_Ea(m)
                enddo
                ! beam axis midpoints (the moment arms of the loads, Solidbody.f90:954)
                do e = 1, VBodies(b)%rbm%nEL
                    xc(:, ke + e) = 0.5d0 * (VBodies(b)%rbm%m_elements(e)%x1(1:3) + &
                                             VBodies(b)%rbm%m_elements(e)%x1(7:9))
                enddo
                km = km + VBodies(b)%v_nelmts
                ke = ke + VBodies(b)%rbm%nEL
            enddo
            call fsil_check(fsil_ibm_interaction_force(h, LBMblks(treenode)%dh, Exyz, Evel, Ea, xc, Eforce, lod, &
                iters, dmax), 'fsil_ibm_interaction_force')
            km = 0; kn = 0
            do i = 1, nb
                b = LBMblks(treenode)%carriedBodies(i)
                VBodies(b)%v_Eforce(1:3, 1:VBodies(b)%v_nelmts) = Eforce(:, km + 1:km + VBodies(b)%v_nelmts)
                VBodies(b)%rbm%lodFlow(1:6 * VBodies(b)%rbm%nND) = lod(kn + 1:kn + 6 * VBodies(b)%rbm%nND)
                km = km + VBodies(b)%v_nelmts
                kn = kn + 6 * VBodies(b)%rbm%nND
            enddo
            dt_solid = LBMblks(treenode)%dh / dble(flow%numsubstep)
            do isubstep = 1, flow%numsubstep
                call Solver(LBMblks(treenode)%carriedBodies, LBMblks(treenode)%blktime, isubstep, &
                            LBMblks(treenode)%dh, dt_solid)
            enddo
        endif
        call fsil_check(fsil_collide_stream_bc(h), 'fsil_collide_stream_bc')
    end subroutine

    ! calculate_macro_quantities_blocks after the step (main.f90:107): the
    ! device's den and uuu into the host arrays the writers read
    subroutine fsil_host_macro()
        integer :: ib
        do ib = 1, size(LBMblks)
            call fsil_check(fsil_macro_download(hnd(ib), LBMblks(ib)%den, LBMblks(ib)%uuu), 'fsil_macro_download')
        enddo
    end subroutine

    ! the device state into fIn before write_continue_blocks (main.f90:118)
    subroutine fsil_host_fin()
        integer :: ib
        do ib = 1, size(LBMblks)
            call fsil_check(fsil_f_download(hnd(ib), LBMblks(ib)%fIn), 'fsil_f_download')
        enddo
    end subroutine

end module fsil_host
