! fsil_bridge.f90 -- ISO_C_BINDING interface to libfsil.so (include/fsilbm3d_amd.h)
! for a Fortran host such as the reference's FSILBM3D program.
!
! Arrays are passed as the host's own Fortran arrays (contiguous, column-major):
! fIn(zDim,yDim,xDim,0:18), den(zDim,yDim,xDim), uuu(zDim,yDim,xDim,1:3),
! v_Exyz(3,M), v_Evel(3,M), v_Ea(M), v_Eforce(3,M), lodFlow(6*nND).
! fsil_check() follows the reference's error convention: print the message
! and stop (e.g. FluidDomain.f90:703-704).
module fsil_bridge
    use, intrinsic :: iso_c_binding
    implicit none
    private
    public :: fsil_init, fsil_finalize, fsil_synchronize, fsil_block_create, fsil_block_destroy, &
              fsil_set_flow, fsil_set_time, fsil_get_volume_force, fsil_f_upload, fsil_f_fill, fsil_f_download, &
              fsil_set_boundary_conditions, fsil_macro_download, fsil_collide_stream_bc, fsil_steps, &
              fsil_ibm_define_bodies, fsil_ibm_set_params, fsil_ibm_interaction_force, fsil_check, &
              fsil_version, fsil_time_collide, fsil_time_steps, fsil_ibm_band_info, fsil_copy_bandwidth, &
              fsil_fem_define_bodies, fsil_fem_set_solver, fsil_fem_update_markers, fsil_fem_substeps, &
              fsil_fem_state, fsil_fem_info, fsil_tree_add_son, fsil_tree_pair_info, &
              fsil_extract_interpolate_layer, fsil_interpolation_father_to_son, fsil_deliver_son_to_father, &
              fsil_tree_steps, fsil_time_tree_steps, fsil_mlp_create, fsil_mlp_destroy, fsil_mlp_get_params, &
              fsil_mlp_set_params, fsil_mlp_forward, fsil_mlp_loss_grad, fsil_mlp_train, fsil_ibm_set_warm_start, &
              fsil_ibm_rounds, fsil_solid_info, fsil_solid_info_static, fsil_set_viscosity, fsil_tau_download, &
              fsil_flow_pack, fsil_write_flow, fsil_write_wait, fsil_write_continue, fsil_fem_checkpoint, fsil_fem_restore, &
              fsil_fsi_step, &
              fsil_ibm_warm_start_data, fsil_ibm_forces, &
              fsil_ibm_report

    interface
        integer(c_int) function fsil_version() bind(C, name='fsil_version')
            import :: c_int
        end function
        type(c_ptr) function fsil_last_error_c() bind(C, name='fsil_last_error')
            import :: c_ptr
        end function
        integer(c_int) function fsil_init(device) bind(C, name='fsil_init')
            import :: c_int
            integer(c_int), value :: device
        end function
        integer(c_int) function fsil_finalize() bind(C, name='fsil_finalize')
            import :: c_int
        end function
        integer(c_int) function fsil_synchronize() bind(C, name='fsil_synchronize')
            import :: c_int
        end function
        integer(c_int) function fsil_block_create(id, dims, geom, bc, iCollidModel, relax, m_collid, &
                                                  m_force, precision, handle) bind(C, name='fsil_block_create')
            import :: c_int, c_double
            integer(c_int), value :: id, iCollidModel, precision
            integer(c_int), intent(in) :: dims(3), bc(6)
            real(c_double), intent(in) :: geom(7), relax(3), m_collid(19,19), m_force(19,19)
            integer(c_int), intent(out) :: handle
        end function
        integer(c_int) function fsil_block_destroy(handle) bind(C, name='fsil_block_destroy')
            import :: c_int
            integer(c_int), value :: handle
        end function
        integer(c_int) function fsil_set_flow(handle, denIn, uvwIn, shearRateIn, velocityKind, volumeForceIn, &
                                              volumeForceAmp, volumeForceFreq, volumeForcePhi) &
                                              bind(C, name='fsil_set_flow')
            import :: c_int, c_double
            integer(c_int), value :: handle, velocityKind
            real(c_double), value :: denIn, volumeForceAmp, volumeForceFreq, volumeForcePhi
            real(c_double), intent(in) :: uvwIn(3), shearRateIn(3), volumeForceIn(3)
        end function
        integer(c_int) function fsil_set_time(handle, blktime) bind(C, name='fsil_set_time')
            import :: c_int, c_double
            integer(c_int), value :: handle
            real(c_double), value :: blktime
        end function
        integer(c_int) function fsil_get_volume_force(handle, vf) bind(C, name='fsil_get_volume_force')
            import :: c_int, c_double
            integer(c_int), value :: handle
            real(c_double), intent(out) :: vf(3)
        end function
        integer(c_int) function fsil_f_upload(handle, fIn) bind(C, name='fsil_f_upload')
            import :: c_int, c_double
            integer(c_int), value :: handle
            real(c_double), intent(in) :: fIn(*)
        end function
        integer(c_int) function fsil_f_fill(handle, f19) bind(C, name='fsil_f_fill')
            import :: c_int, c_double
            integer(c_int), value :: handle
            real(c_double), intent(in) :: f19(19)
        end function
        integer(c_int) function fsil_f_download(handle, fIn) bind(C, name='fsil_f_download')
            import :: c_int, c_double
            integer(c_int), value :: handle
            real(c_double), intent(out) :: fIn(*)
        end function
        integer(c_int) function fsil_set_boundary_conditions(handle) bind(C, name='fsil_set_boundary_conditions')
            import :: c_int
            integer(c_int), value :: handle
        end function
        integer(c_int) function fsil_macro_download(handle, den, uuu) bind(C, name='fsil_macro_download')
            import :: c_int, c_double
            integer(c_int), value :: handle
            real(c_double), intent(out) :: den(*), uuu(*)
        end function
        integer(c_int) function fsil_collide_stream_bc(handle) bind(C, name='fsil_collide_stream_bc')
            import :: c_int
            integer(c_int), value :: handle
        end function
        integer(c_int) function fsil_steps(handle, nsteps, t0, dt) bind(C, name='fsil_steps')
            import :: c_int, c_double
            integer(c_int), value :: handle, nsteps
            real(c_double), value :: t0, dt
        end function
        integer(c_int) function fsil_ibm_define_bodies(handle, nbodies, nmarkers, nnodes, nelems, marker_elem, &
                                                       elem_nodes) bind(C, name='fsil_ibm_define_bodies')
            import :: c_int
            integer(c_int), value :: handle, nbodies
            integer(c_int), intent(in) :: nmarkers(*), nnodes(*), nelems(*), marker_elem(*), elem_nodes(*)
        end function
        integer(c_int) function fsil_ibm_set_params(handle, ntolLBM, dtolLBM, Uref, denIn) &
                                                    bind(C, name='fsil_ibm_set_params')
            import :: c_int, c_double
            integer(c_int), value :: handle, ntolLBM
            real(c_double), value :: dtolLBM, Uref, denIn
        end function
        ! every array and output is optional: an absent one arrives as a null
        ! pointer (Fortran 2018), so a host with its bodies on the device beam
        ! solver omits the markers (fsil_fem_update_markers has formed them),
        ! and a host omits the outputs it does not want back (no host wait;
        ! fsil_ibm_report and fsil_ibm_forces read them later), e.g.
        !   rc = fsil_ibm_interaction_force(h, dt, iters=it, dmax=dm)
        integer(c_int) function fsil_ibm_interaction_force(handle, dt, Exyz, Evel, Ea, elem_xc, Eforce, lodFlow, &
                                                           iters, dmax) bind(C, name='fsil_ibm_interaction_force')
            import :: c_int, c_double
            integer(c_int), value :: handle
            real(c_double), value :: dt
            real(c_double), intent(in), optional :: Exyz(*), Evel(*), Ea(*), elem_xc(*)
            real(c_double), intent(out), optional :: Eforce(*), lodFlow(*)
            integer(c_int), intent(out), optional :: iters
            real(c_double), intent(out), optional :: dmax
        end function
        integer(c_int) function fsil_time_collide(handle, nlaunch, avg_ms) bind(C, name='fsil_time_collide')
            import :: c_int, c_double
            integer(c_int), value :: handle, nlaunch
            real(c_double), intent(out) :: avg_ms
        end function
        integer(c_int) function fsil_time_steps(handle, nsteps, t0, dt, total_ms) bind(C, name='fsil_time_steps')
            import :: c_int, c_double
            integer(c_int), value :: handle, nsteps
            real(c_double), value :: t0, dt
            real(c_double), intent(out) :: total_ms
        end function
        integer(c_int) function fsil_ibm_forces(handle, Eforce, lodFlow) bind(C, name='fsil_ibm_forces')
            import :: c_int, c_double
            integer(c_int), value :: handle
            real(c_double), intent(out) :: Eforce(*), lodFlow(*)
        end function
        integer(c_int) function fsil_ibm_report(handle, iters, dmax) bind(C, name='fsil_ibm_report')
            import :: c_int, c_double
            integer(c_int), value :: handle
            integer(c_int), intent(out) :: iters
            real(c_double), intent(out) :: dmax
        end function
        integer(c_int) function fsil_ibm_band_info(handle, nband, bbox) bind(C, name='fsil_ibm_band_info')
            import :: c_int
            integer(c_int), value :: handle
            integer(c_int), intent(out) :: nband, bbox(6)
        end function
        integer(c_int) function fsil_copy_bandwidth(nbytes, reps, gbs) bind(C, name='fsil_copy_bandwidth')
            import :: c_int, c_long_long, c_double
            integer(c_long_long), value :: nbytes
            integer(c_int), value :: reps
            real(c_double), intent(out) :: gbs
        end function
        ! beam bodies on the device (SolidSolver.f90, Solidbody.f90)
        integer(c_int) function fsil_fem_define_bodies(handle, nbodies, nND, nEL, xyz8, elems, cons, iBodyModel, &
                                                       isMotionGiven, mat, kin, Lref, Uref, denIn, alpha) &
                                                       bind(C, name='fsil_fem_define_bodies')
            import :: c_int, c_double
            integer(c_int), value :: handle, nbodies
            integer(c_int), intent(in) :: nND(*), nEL(*), elems(*), cons(*), iBodyModel(*), isMotionGiven(*)
            real(c_double), intent(in) :: xyz8(*), mat(*), kin(*)
            real(c_double), value :: Lref, Uref, denIn, alpha
        end function
        integer(c_int) function fsil_fem_set_solver(handle, dampK, dampM, GeoGamma, NewmarkGamma, NewmarkBeta, &
                                                    dtolFEM, ntolFEM) bind(C, name='fsil_fem_set_solver')
            import :: c_int, c_double
            integer(c_int), value :: handle, ntolFEM
            real(c_double), value :: dampK, dampM, GeoGamma, NewmarkGamma, NewmarkBeta, dtolFEM
        end function
        integer(c_int) function fsil_fem_update_markers(handle) bind(C, name='fsil_fem_update_markers')
            import :: c_int
            integer(c_int), value :: handle
        end function
        integer(c_int) function fsil_fem_substeps(handle, time, numsubstep, deltat) bind(C, name='fsil_fem_substeps')
            import :: c_int, c_double
            integer(c_int), value :: handle, numsubstep
            real(c_double), value :: time, deltat
        end function
        integer(c_int) function fsil_fem_state(handle, body, pos, vel, acc, dsp) bind(C, name='fsil_fem_state')
            import :: c_int, c_double
            integer(c_int), value :: handle, body
            real(c_double), intent(out) :: pos(*), vel(*), acc(*), dsp(*)
        end function
        integer(c_int) function fsil_fem_info(handle, body, newton_iters, cg_iters, dnorm) bind(C, name='fsil_fem_info')
            import :: c_int, c_double
            integer(c_int), value :: handle, body
            integer(c_int), intent(out) :: newton_iters, cg_iters
            real(c_double), intent(out) :: dnorm
        end function
        ! block tree (LBMBlockComm.f90)
        integer(c_int) function fsil_tree_add_son(father, son, interpolateScheme, index) &
                                                  bind(C, name='fsil_tree_add_son')
            import :: c_int
            integer(c_int), value :: father, son, interpolateScheme
            integer(c_int), intent(out) :: index
        end function
        integer(c_int) function fsil_tree_pair_info(father, index, comm) bind(C, name='fsil_tree_pair_info')
            import :: c_int
            integer(c_int), value :: father, index
            integer(c_int), intent(out) :: comm(36)
        end function
        integer(c_int) function fsil_extract_interpolate_layer(father, level) &
                                                               bind(C, name='fsil_extract_interpolate_layer')
            import :: c_int
            integer(c_int), value :: father, level
        end function
        integer(c_int) function fsil_interpolation_father_to_son(father, index, n_timeStep) &
                                                                 bind(C, name='fsil_interpolation_father_to_son')
            import :: c_int
            integer(c_int), value :: father, index, n_timeStep
        end function
        integer(c_int) function fsil_deliver_son_to_father(father, index) bind(C, name='fsil_deliver_son_to_father')
            import :: c_int
            integer(c_int), value :: father, index
        end function
        integer(c_int) function fsil_tree_steps(root, nsteps, t0, dt) bind(C, name='fsil_tree_steps')
            import :: c_int, c_double
            integer(c_int), value :: root, nsteps
            real(c_double), value :: t0, dt
        end function
        integer(c_int) function fsil_time_tree_steps(root, nsteps, t0, dt, total_ms) &
                                                     bind(C, name='fsil_time_tree_steps')
            import :: c_int, c_double
            integer(c_int), value :: root, nsteps
            real(c_double), value :: t0, dt
            real(c_double), intent(out) :: total_ms
        end function
        ! learned warm start of the direct-forcing loop (no reference counterpart)
        integer(c_int) function fsil_mlp_create(seed, handle) bind(C, name='fsil_mlp_create')
            import :: c_int
            integer(c_int), value :: seed
            integer(c_int), intent(out) :: handle
        end function
        integer(c_int) function fsil_mlp_destroy(handle) bind(C, name='fsil_mlp_destroy')
            import :: c_int
            integer(c_int), value :: handle
        end function
        integer(c_int) function fsil_mlp_get_params(handle, theta) bind(C, name='fsil_mlp_get_params')
            import :: c_int, c_float
            integer(c_int), value :: handle
            real(c_float), intent(out) :: theta(*)
        end function
        integer(c_int) function fsil_mlp_set_params(handle, theta) bind(C, name='fsil_mlp_set_params')
            import :: c_int, c_float
            integer(c_int), value :: handle
            real(c_float), intent(in) :: theta(*)
        end function
        integer(c_int) function fsil_mlp_forward(handle, X, n, Y, bf16) bind(C, name='fsil_mlp_forward')
            import :: c_int, c_float
            integer(c_int), value :: handle, n, bf16
            real(c_float), intent(in) :: X(*)
            real(c_float), intent(out) :: Y(*)
        end function
        integer(c_int) function fsil_mlp_loss_grad(handle, X, Y, n, loss, grad) bind(C, name='fsil_mlp_loss_grad')
            import :: c_int, c_float, c_double
            integer(c_int), value :: handle, n
            real(c_float), intent(in) :: X(*), Y(*)
            real(c_double), intent(out) :: loss
            real(c_float), intent(out) :: grad(*)
        end function
        integer(c_int) function fsil_mlp_train(handle, X, Y, n, epochs, batch, lr, seed, loss_before, loss_after) &
                                               bind(C, name='fsil_mlp_train')
            import :: c_int, c_float, c_double
            integer(c_int), value :: handle, n, epochs, batch, seed
            real(c_double), value :: lr
            real(c_float), intent(in) :: X(*), Y(*)
            real(c_double), intent(out) :: loss_before, loss_after
        end function
        integer(c_int) function fsil_ibm_set_warm_start(handle, mlp, record, skip) bind(C, name='fsil_ibm_set_warm_start')
            import :: c_int
            integer(c_int), value :: handle, mlp, record, skip
        end function
        integer(c_int) function fsil_solid_info(handle, body, refs, rows) bind(C, name='fsil_solid_info')
            import :: c_int, c_double
            integer(c_int), value :: handle, body
            real(c_double), intent(in) :: refs(6)
            real(c_double), intent(out) :: rows(78)
        end function
        integer(c_int) function fsil_solid_info_static(handle, body, refs, rows) &
                bind(C, name='fsil_solid_info_static')
            import :: c_int, c_double
            integer(c_int), value :: handle, body
            real(c_double), intent(in) :: refs(6)
            real(c_double), intent(out) :: rows(78)
        end function
        integer(c_int) function fsil_set_viscosity(handle, nu) bind(C, name='fsil_set_viscosity')
            import :: c_int, c_double
            integer(c_int), value :: handle
            real(c_double), value :: nu
        end function
        integer(c_int) function fsil_tau_download(handle, tau_all) bind(C, name='fsil_tau_download')
            import :: c_int, c_double
            integer(c_int), value :: handle
            real(c_double), intent(out) :: tau_all(*)
        end function
        integer(c_int) function fsil_flow_pack(handle, offsetOutput, denIn, Uref, out) bind(C, name='fsil_flow_pack')
            import :: c_int, c_double, c_float
            integer(c_int), value :: handle, offsetOutput
            real(c_double), value :: denIn, Uref
            real(c_float), intent(out) :: out(*)
        end function
        integer(c_int) function fsil_write_flow(handle, offsetOutput, denIn, Uref, path) bind(C, name='fsil_write_flow')
            import :: c_int, c_double, c_char
            integer(c_int), value :: handle, offsetOutput
            real(c_double), value :: denIn, Uref
            character(kind=c_char), intent(in) :: path(*)
        end function
        integer(c_int) function fsil_write_wait() bind(C, name='fsil_write_wait')
            import :: c_int
        end function
        integer(c_int) function fsil_write_continue(nblocks, handles, step, time, path) &
                bind(C, name='fsil_write_continue')
            import :: c_int, c_double, c_char
            integer(c_int), value :: nblocks, step
            integer(c_int), intent(in) :: handles(*)
            real(c_double), value :: time
            character(kind=c_char), intent(in) :: path(*)
        end function
        integer(c_int) function fsil_fem_checkpoint(handle, buf, nbytes) bind(C, name='fsil_fem_checkpoint')
            import :: c_int, c_ptr, c_long_long
            integer(c_int), value :: handle
            type(c_ptr), value :: buf
            integer(c_long_long), intent(inout) :: nbytes
        end function
        integer(c_int) function fsil_fem_restore(handle, buf, nbytes) bind(C, name='fsil_fem_restore')
            import :: c_int, c_ptr, c_long_long
            integer(c_int), value :: handle
            type(c_ptr), value :: buf
            integer(c_long_long), value :: nbytes
        end function
        integer(c_int) function fsil_fsi_step(handle, time, numsubstep, iters, dmax) bind(C, name='fsil_fsi_step')
            import :: c_int, c_double
            integer(c_int), value :: handle, numsubstep
            real(c_double), value :: time
            integer(c_int), intent(out) :: iters
            real(c_double), intent(out) :: dmax
        end function
        integer(c_int) function fsil_ibm_rounds(handle, rounds) bind(C, name='fsil_ibm_rounds')
            import :: c_int
            integer(c_int), value :: handle
            integer(c_int), intent(out) :: rounds
        end function
        integer(c_int) function fsil_ibm_warm_start_data(handle, X, Y, info) bind(C, name='fsil_ibm_warm_start_data')
            import :: c_int, c_float
            integer(c_int), value :: handle
            real(c_float), intent(out) :: X(*), Y(*)
            integer(c_int), intent(out) :: info(2)
        end function
    end interface

contains

    subroutine fsil_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        character(len=1024) :: text
        integer :: i
        if (rc == 0) return
        call c_f_pointer(fsil_last_error_c(), msg, [1024])
        text = ' '
        do i = 1, 1024
            if (msg(i) == c_null_char) exit
            text(i:i) = msg(i)
        enddo
        write(*,*) 'fsil error in ', where, ': ', trim(text)
        stop
    end subroutine fsil_check

end module fsil_bridge
