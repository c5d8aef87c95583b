/*
 * fsilbm3d_amd.h -- C-ABI of the MI355X-native FSILBM3D fluid/coupling hot
 * path (libfsil.so, gfx950).
 *
 * Every entry point returns 0 on success and nonzero on error; the message is
 * then available from fsil_last_error() (the reference's convention is to
 * `stop` with a message, which the Fortran binding in include/fsil_bridge.f90
 * does on a nonzero return).  Arguments are plain pointers and sizes; arrays
 * are the reference's host arrays in Fortran memory order.  The device owns
 * the distribution functions of each block (A-B buffers, SoA by direction,
 * z fastest) between calls; the host sees them only through
 * fsil_f_download / fsil_macro_download.
 *
 * Reference citations are to gaoak/FSILBM3D src/.
 */
#ifndef FSILBM3D_AMD_H
#define FSILBM3D_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* Library version (1). */
int fsil_version(void);

/* Message of the last failed call on this thread. */
const char *fsil_last_error(void);

/* Select the HIP device; must precede fsil_block_create.  No reference
 * counterpart (the reference is a CPU program; cf. omp_set_num_threads at
 * main.f90:36). */
int fsil_init(int device);

/* Destroy every block and release device memory. */
int fsil_finalize(void);

/* Wait for all device work. */
int fsil_synchronize(void);

/* Create one fluid block: allocate_fluid_ + the relaxation part of
 * initialise_ (FluidDomain.f90:378-431, 433-522).
 *   dims[3]   xDim, yDim, zDim                     (FluidDomain.f90:73)
 *   geom[7]   dh, xmin, ymin, zmin, xmax, ymax, zmax (FluidDomain.f90:75,93-104)
 *   bc[6]     BndConds x-,x+,y-,y+,z-,z+: 0,1,101-104,201-204,301,302
 *   iCollidModel 1 SRT, 2 TRT, 3 MRT              (FluidDomain.f90:1225-1238),
 *             11 Smagorinsky, 14 WALE, 15 Vreman LES (FluidDomain.f90:1239-1258, 1265-1507;
 *             FP64 form, single blocks; 14 and 15 need fsil_set_viscosity)
 *   relax[3]  tau, Omega, Omega2                    (FluidDomain.f90:449-464)
 *   m_collid, m_force  19x19 Fortran-order M_COLLID, M_FORCE for MRT, else NULL
 *   precision 0: FP64 plain; 1: FP32 storage of f - w (SRT; every boundary code but 102/203/204)
 *   handle    out: block handle */
int fsil_block_create(int id, const int *dims, const double *geom, const int *bc, int iCollidModel,
                      const double *relax, const double *m_collid, const double *m_force,
                      int precision, int *handle);

/* Free a block. */
int fsil_block_destroy(int handle);

/* Flow conditions the block's boundaries and volume force read from the
 * reference's `flow` global (FlowCondition.f90:11-26): denIn, uvwIn[3],
 * shearRateIn[3], velocityKind, volumeForceIn[3], volumeForceAmp/Freq/Phi. */
int fsil_set_flow(int handle, double denIn, const double *uvwIn, const double *shearRateIn,
                  int velocityKind, const double *volumeForceIn, double volumeForceAmp,
                  double volumeForceFreq, double volumeForcePhi);

/* Set LBMblks%blktime and evaluate update_volume_force_ at it
 * (main.f90:97, FluidDomain.f90:1174-1180). */
int fsil_set_time(int handle, double blktime);

/* flow%nu (Solidbody.f90:282), the molecular viscosity the WALE and Vreman
 * models add their eddy viscosity to (FluidDomain.f90:1421, 1504). */
int fsil_set_viscosity(int handle, double nu);

/* The per-cell relaxation time tau_all(z,y,x) (FluidDomain.f90:35, 454-455):
 * written by the LES models each step; tau everywhere for models 1-3. */
int fsil_tau_download(int handle, double *tau_all);

/* The volume force of the current block time (vf[3]). */
int fsil_get_volume_force(int handle, double *vf);

/* Copy fIn(z,y,x,0:18) (pre-collision, post-boundary state) to / from the
 * device.  Replaces nothing: the reference's fIn lives in host memory. */
int fsil_f_upload(int handle, const double *fIn);
/* Every cell set to the same 19 populations f19[q]: the uniform state of
 * initialise_flow (FluidDomain.f90:524-544) without a host copy of the field. */
int fsil_f_fill(int handle, const double *f19);
int fsil_f_download(int handle, double *fIn);

/* set_boundary_conditions_ on the stored state as called once at start-up
 * (main.f90:63; FluidDomain.f90:616-1126). */
int fsil_set_boundary_conditions(int handle);

/* calculate_macro_quantities_ of the stored state (FluidDomain.f90:1128-1145,
 * as main.f90:107 calls it): den(z,y,x), uuu(z,y,x,1:3). */
int fsil_macro_download(int handle, double *den, double *uuu);

/* ---------------------------------------------------------------------------
 * Output and restart (FluidDomain.f90:268-285, 1628-1736, 1770-1789;
 * SolidSolver.f90:1669-1692).  The host keeps file naming and cadence. */

/* The instantaneous DatFlow fields of write_flow_ (outputtype 1) computed on
 * the device from the stored state: out[4][nx*ny*nz] real(4) p = Cs2 (den -
 * denIn), u, v, w = uuu / Uref over the output window (offsetOutput cells left
 * out on every face), z fastest. */
int fsil_flow_pack(int handle, int offsetOutput, double denIn, double Uref, float *out);

/* write_flow_ without the fork: packs on the device, copies into pinned host
 * memory on the block's stream and returns; a host thread writes `path` in the
 * reference's stream layout (nx, ny, nz, ID int32; xmin, ymin, zmin, dh real8
 * of the window; p, u, v, w real(4)) once the copy completes. */
int fsil_write_flow(int handle, int offsetOutput, double denIn, double Uref, const char *path);

/* Wait for every pending fsil_write_flow (the reference's mywait loop before
 * the next write, FluidDomain.f90:355-357); reports the first write error. */
int fsil_write_wait(void);

/* write_continue_blocks (FluidDomain.f90:268-285, 1770-1776) from the device
 * states of nblocks blocks: nblocks, step (int32), time / Tref (real8), then
 * per block xmin, ymin, zmin, dh (real8), xDim, yDim, zDim (int32), fIn. */
int fsil_write_continue(int nblocks, const int *handles, int step, double time, const char *path);

/* Checkpoint of the block's beam bodies (Beam_write_solid_temp / read, which
 * the reference declares at SolidSolver.f90:1669-1692): element records, then
 * pos, dsp, vel, acc, lodInte, lodExte, lodEffe, lodFlow, vBC.  With buf NULL
 * *bytes receives the size; restore needs the same bodies defined. */
int fsil_fem_checkpoint(int handle, void *buf, long long *bytes);
int fsil_fem_restore(int handle, const void *buf, long long bytes);

/* One fluid update of a block: collision_ (with the immersed-boundary
 * velocity/force of the last fsil_ibm_interaction_force on its band cells,
 * the uniform volume force elsewhere), halfwayBCset_, streaming_ and
 * set_boundary_conditions_ (LBMBlockComm.f90:293-304; FluidDomain.f90:567-614,
 * 616-1126, 1208-1263, 1514-1625), fused into one pass plus boundary faces. */
int fsil_collide_stream_bc(int handle);

/* nsteps fluid-only steps at block times t0+dt*(s+1) (volume force updated
 * each step); the body of main.f90:93-147 for a block carrying no bodies. */
int fsil_steps(int handle, int nsteps, double t0, double dt);

/* Timing helpers (HIP events on the block's stream): average duration of the
 * fused collide+stream kernel over nlaunch launches; total of nsteps steps. */
int fsil_time_collide(int handle, int nlaunch, double *avg_ms);
int fsil_time_steps(int handle, int nsteps, double t0, double dt, double *total_ms);

/* Immersed bodies carried by the block (FindCarrierFluidBlock,
 * FluidDomain.f90:1998-2017; PlateBuild_, Solidbody.f90:1051-1071):
 * nmarkers/nnodes/nelems per body; marker_elem[M] element (0-based, within
 * its body) of each marker (vtor); elem_nodes[2*sum(nelems)] node0/node1
 * (0-based, within the body) of each element. */
int fsil_ibm_define_bodies(int handle, int nbodies, const int *nmarkers, const int *nnodes,
                           const int *nelems, const int *marker_elem, const int *elem_nodes);

/* ntolLBM, dtolLBM, Uref, denIn of the direct-forcing loop
 * (set_solidbody_parameters, Solidbody.f90:328-348). */
int fsil_ibm_set_params(int handle, int ntolLBM, double dtolLBM, double Uref, double denIn);

/* FSInteraction_force after UpdatePosVelArea_ (Solidbody.f90:589-602,
 * 760-1049): discrete-delta stencils, the multi-direct-forcing loop on the
 * band macros of the stored state, force spreading and the beam loads.
 *   Exyz[3M], Evel[3M], Ea[M]  markers of all bodies, carried order
 *   elem_xc[3*nelems]          beam axis midpoints (moment arm origin)
 *   Eforce[3M] out             v_Eforce;  lodFlow[6*sum(nnodes)] out
 *   iters, dmax out            iterations and final residual
 * With Eforce, lodFlow, iters and dmax all NULL the call only queues its work
 * on the block's stream and returns (the fluid update and beam sub-steps that
 * follow are ordered after it on the device); its errors are then reported
 * by the next coupling call or fsil_ibm_forces. */
int fsil_ibm_interaction_force(int handle, double dt, const double *Exyz, const double *Evel,
                               const double *Ea, const double *elem_xc, double *Eforce,
                               double *lodFlow, int *iters, double *dmax);

/* v_Eforce and lodFlow of the last fsil_ibm_interaction_force (either may be
 * NULL): the host reads of the bodies' marker forces and beam loads that the
 * reference's writers make (Solidbody.f90:920-979 leaves them in the body
 * records); the coupling call itself need not copy them out every step. */
int fsil_ibm_forces(int handle, double *Eforce, double *lodFlow);

/* Iteration count and final residual of the last fsil_ibm_interaction_force
 * (either pointer may be NULL), waiting for that call if it was made without
 * read-backs, with its errors reported here: the values calculate_interaction_force
 * (Solidbody.f90:869-918) ends its loop with.  Lets a caller queue the beam
 * sub-steps behind a coupling call on the device before it waits for the call. */
int fsil_ibm_report(int handle, int *iters, double *dmax);

/* Plate bodies on the device beam solver, carried by the block (read_solid_files,
 * Beam_ReadBuild, Beam_Initialise, calculate_angle_material with isKB = 1:
 * Solidbody.f90:73-216, SolidSolver.f90:1270-1446, 1533-1616).  Per body b:
 *   nND[b], nEL[b]; xyz8: nND rows "x y z Lspan Rspan dirX dirY dirZ" (plate.dat POINT);
 *   elems: nEL rows "i j Nspan" (1-based nodes); cons: nND rows of 6 (CONSTRAINT);
 *   iBodyModel[b] 1 rigid / 2 elastic; isMotionGiven[6b..6b+5];
 *   mat[4b..]: denR, KB, KS, psR;  kin[22b..]: Freq, XYZo[3], XYZAmpl[3], XYZPhi[3] (deg),
 *   AoAo[3], AoAAmpl[3], AoAPhi[3] (deg), initXYZVel[3];  Lref, Uref, denIn, IBPenaltyAlpha.
 * Also defines the bodies' immersed-boundary markers (as fsil_ibm_define_bodies). */
int fsil_fem_define_bodies(int handle, int nbodies, const int *nND, const int *nEL, const double *xyz8,
                           const int *elems, const int *cons, const int *iBodyModel, const int *isMotionGiven,
                           const double *mat, const double *kin, double Lref, double Uref, double denIn,
                           double alpha);

/* Set_SolidSolver_Params (SolidSolver.f90:1249-1264). */
int fsil_fem_set_solver(int handle, double dampK, double dampM, double GeoGamma, double NewmarkGamma,
                        double NewmarkBeta, double dtolFEM, int ntolFEM);

/* PlateUpdatePosVelArea_ for every FEM body on the device (Solidbody.f90:604-646);
 * fsil_ibm_interaction_force with Exyz == NULL then uses these markers. */
int fsil_fem_update_markers(int handle);

/* numsubstep calls of Beam_structure for every body (LBMBlockComm.f90:333-335,
 * SolidSolver.f90:1820-2359), sub-step deltat/numsubstep, loads = the lodFlow of
 * the last fsil_ibm_interaction_force. */
int fsil_fem_substeps(int handle, double time, int numsubstep, double deltat);

/* Beam_write_solid_info for body b (SolidSolver.f90:1729-1799): the values of
 * one row of each DatInfo/Group<g>_<file>.dat the reference appends at info
 * cadence, computed on the device from the last beam sub-steps and coupling
 * call, every sum in the reference's order.
 *   refs[6]   Lref, Uref, Aref, Fref, Pref, Eref (Solidbody.f90:276-280)
 *   rows[78]  firstNode[15] lastNode[15] centerNode[15] nodeAverage[12]
 *             forces[6] power[7] energy[8], columns as the reference writes them
 * The host keeps the file writing (formats, ZONE lines). */
int fsil_solid_info(int handle, int body, const double *refs, double *rows);

/* fsil_solid_info without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  Its
 * kernel works in a static device array of the library (one per device, set
 * up by fsil_init) that holds bodies of at most 256 elements; a larger body
 * is an error.  Calls that overlap on one device share that array and race. */
int fsil_solid_info_static(int handle, int body, const double *refs, double *rows);

/* One time step of a block whose bodies are on the device beam solver: the
 * body of main.f90:95-107 for that block (tree_collision_streaming_IBM_FEM +
 * IBM_FEM, LBMBlockComm.f90:279-338) in one call: block time and volume force,
 * the device markers, the coupling call, numsubstep beam sub-steps of dh /
 * numsubstep, and the fused collide/stream/boundary update.  iters and dmax
 * (either may be NULL) receive the forcing loop's count and residual. */
int fsil_fsi_step(int handle, double time, int numsubstep, int *iters, double *dmax);

/* Node state of body b (6*nND each, node-major), and the last sub-step's
 * Newton iterations, conjugate-gradient iterations and displacement norm. */
int fsil_fem_state(int handle, int body, double *pos, double *vel, double *acc, double *dsp);
int fsil_fem_info(int handle, int body, int *newton_iters, int *cg_iters, double *dnorm);

/* Bandwidth of a plain device copy of nbytes (read + write), reps launches:
 * the calibration for HBM counters and the achievable-bandwidth context of the
 * roofline.  No reference counterpart (measurement helper). */
int fsil_copy_bandwidth(long long nbytes, int reps, double *gbs);

/* Band size and bounding box (x0,y0,z0,nx,ny,nz) of the last coupling call. */
int fsil_ibm_band_info(int handle, int *nband, int *bbox);

/* ---------------------------------------------------------------------------
 * Block tree: refined son blocks (dh = father dh / 2) whose BCfluid (0) faces
 * are fed from the father (LBMBlockComm.f90).  Tree blocks use the FP64 plain
 * form; a son shares its father's stream. */

/* Make `son` a son of `father`: check_blocks_params (LBMBlockComm.f90:508-544),
 * the CommPair of build_blocks_comunication (:33-97) and the father planes of
 * allocate_fIn_tau (:201-246).  interpolateScheme 2: 3rd/4th-order plane
 * interpolation, otherwise linear (FlowCondition%interpolateScheme).
 * *index receives the son's position among the father's sons (0-based). */
int fsil_tree_add_son(int father, int son, int interpolateScheme, int *index);

/* CommPair of son `index`: comm[0..35] = sds[6], s[6], f[6], si[6], fi[6]
 * (1-based, as LBMBlockComm.f90:11-18), xDimS yDimS zDimS, xDimF yDimF zDimF. */
int fsil_tree_pair_info(int father, int index, int *comm);

/* extract_interpolate_layer (LBMBlockComm.f90:340-505) for every son of
 * `father`: level 1 before the father's collision, level 2 after its boundary
 * faces (level 2 also averages level 1 with it). */
int fsil_extract_interpolate_layer(int father, int level);

/* interpolation_father_to_son (LBMBlockComm.f90:655-806) after the son's
 * sub-step n_timeStep (0 or 1). */
int fsil_interpolation_father_to_son(int father, int index, int n_timeStep);

/* deliver_son_to_father (LBMBlockComm.f90:546-653). */
int fsil_deliver_son_to_father(int father, int index);

/* nsteps root steps of tree_collision_streaming_IBM_FEM (LBMBlockComm.f90:279-318)
 * for a tree whose blocks carry no bodies, time accumulated from t0 by dt and
 * given to every block of the tree each step (main.f90:95-97). */
int fsil_tree_steps(int root, int nsteps, double t0, double dt);

/* fsil_tree_steps bracketed by device events on the tree's stream (ms). */
int fsil_time_tree_steps(int root, int nsteps, double t0, double dt, double *total_ms);

/* ---------------------------------------------------------------------------
 * Learned warm start of the multi-direct-forcing loop.  No counterpart in the
 * reference: it extends calculate_interaction_force (Solidbody.f90:869-918)
 * with a per-marker network 12 -> 64 -> 64 -> 64 -> 3 (tanh hidden layers).
 * The loop is linear in the accumulated marker force, so a guess of the force
 * E_k that `skip` plain rounds reach, applied as one correction of the band
 * velocity, puts the loop in the state of round `skip`; the loop then goes on
 * from iteration skip+1 to the same cap ntolLBM and tolerance dtolLBM.  The
 * network predicts E_k - Eprev (Eprev: the marker's final force of the
 * previous call) over Uref |Ea|.  Parameters: 9347 floats, W1[64][12] b1
 * W2[64][64] b2 W3[64][64] b3 W4[3][64] b4. */

/* A network with Glorot-uniform weights from the splitmix64 stream of seed. */
int fsil_mlp_create(int seed, int *handle);
int fsil_mlp_destroy(int handle);
int fsil_mlp_get_params(int handle, float *theta);
/* Also resets the Adam moments. */
int fsil_mlp_set_params(int handle, const float *theta);
/* Y[n][3] = net(X[n][12]); bf16 != 0: BF16 matrix-core inference (FP32
 * accumulation), else the FP32 training forward. */
int fsil_mlp_forward(int handle, const float *X, int n, float *Y, int bf16);
/* Mean squared error over the n x 3 outputs and its gradient (FP32). */
int fsil_mlp_loss_grad(int handle, const float *X, const float *Y, int n, double *loss, float *grad);
/* Adam on mini-batches of `batch` rows, `epochs` passes in a splitmix64
 * shuffled order; full-data loss before and after (either may be NULL). */
int fsil_mlp_train(int handle, const float *X, const float *Y, int n, int epochs, int batch, double lr, int seed,
                   double *loss_before, double *loss_after);

/* Use network `mlp` (0: none) to guess the force after `skip` rounds
 * (0 <= skip < ntolLBM) in every coupling call of block `handle` that has a
 * previous call; record != 0 computes the network's features and, in plain
 * calls, its training targets (E_skip - Eprev) / (Uref |Ea|).  A guessed call
 * whose first residual exceeds twice the residual the previous call had at
 * iteration skip+1 restarts from the zero guess and runs the plain loop. */
int fsil_ibm_set_warm_start(int handle, int mlp, int record, int skip);

/* Forcing rounds the last fsil_ibm_interaction_force ran: interpolation and
 * correction rounds, plus one for the correction that applies a learned
 * guess (a restarted call counts every round it ran). */
int fsil_ibm_rounds(int handle, int *rounds);

/* Features X[M][12] (interpolated velocity / Uref, marker velocity / Uref,
 * previous force / (Uref |Ea|), normal) and targets Y[M][3] ((E_skip - Eprev)
 * / (Uref |Ea|), plain calls) of the last coupling call; info[0] = guess used,
 * info[1] = fell back to zero. */
int fsil_ibm_warm_start_data(int handle, float *X, float *Y, int *info);

#ifdef __cplusplus
}
#endif
#endif /* FSILBM3D_AMD_H */
