/* C ABI of the MI355X DG right-hand-side library (libmgdg.so).
 *
 * Every entry point takes device pointers (HBM, caller-allocated), sizes and
 * a hipStream_t (passed as void*).  Nothing allocates except where stated,
 * nothing synchronises the stream, and nothing throws across the ABI: a
 * nonzero mg_status is returned and mg_last_error() gives the message.
 *
 * Layouts (N = nelements * np nodes, F = dim + 2 + nspecies fields):
 *   conserved state q         [F][N]    field order mass, energy,
 *                                       momentum[dim], species[nspecies]
 *                                       (mirgecom/fluid.py:311-333)
 *   gradient g                [(F+1)][dim][N]   grad of each conserved field,
 *                                       then grad T
 *   boundary exterior state   [F][nbnd_nodes], temperature [nbnd_nodes]
 *
 * The reference's operator interface is Python-level; the Python host layer
 * (mirgecom_amd.euler / mirgecom_amd.navierstokes) binds these entry points
 * with ctypes behind the reference's signatures.
 */
#ifndef MGDG_H
#define MGDG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t mg_status;
#define MG_OK 0
#define MG_ERR_ARG 1
#define MG_ERR_UNSUPPORTED 2
#define MG_ERR_HIP 3

/* Discretization: reference-element operators, affine geometry and face
 * connectivity, all resident in HBM (built once on the host; see
 * create_discretization_collection, mirgecom/discretization.py:41-83). */
typedef struct mg_mesh {
  int32_t dim, order, np, nfp, nfaces, pad0;
  int64_t nelements;
  int64_t nbnd_nodes;
  const double* rx;         /* [K][dim][dim]  dr_a/dx_i                    */
  const double* normals;    /* [K][nfaces][dim] outward unit normals       */
  const double* fscale;     /* [K][nfaces]  face Jacobian / volume Jacobian */
  const int32_t* vmap_p;    /* [K][nfaces][nfp] neighbour node, or -(b+1)  */
  const double* weak_diff;  /* [dim][np][np]  M^-1 D_r^T M                 */
  const double* lift;       /* [np][nfaces*nfp]  M^-1 E                    */
  const int32_t* face_mask; /* [nfaces][nfp] volume node of each face node */
  const double* workspace;  /* packed operators, set by mg_mesh_prepare     */
  /* Optional (NULL = natural order): elem_rank[e] is element e's position
   * in a locality-preserving traversal (a permutation of 0..K-1, e.g. a
   * space-filling curve through the element centroids).  Workgroups then
   * take their element blocks in that order, so that the blocks one XCD
   * runs together are mesh neighbours and the face gathers hit its L2.
   * Results do not depend on it; a rank array that is not a permutation
   * falls back to the natural order. */
  const int32_t* elem_rank;
  const int32_t* block_order; /* visiting-order tables, set by mg_mesh_prepare */
} mg_mesh;

/* Bytes of the mesh workspace: the packed operators, and the visiting-order
 * tables when elem_rank is set (-1 if dim/order are unsupported). */
int64_t mg_mesh_workspace_bytes(const mg_mesh* mesh);

/* Pack weak_diff and lift into the caller-allocated workspace in the layout
 * the contraction phase reads (rows of each wave contiguous, read with
 * wave-uniform loads) and store the pointer in mesh->workspace; with
 * elem_rank set, also build the block visiting-order tables (one per
 * elements-per-workgroup count the kernels use) and store them in
 * mesh->block_order.  Once per mesh; asynchronous on the stream.  There is
 * no reference counterpart: it is the device-side part of building the
 * discretization (mirgecom/discretization.py:41-83). */
mg_status mg_mesh_prepare(mg_mesh* mesh, double* workspace, void* stream);

/* Gas model: equation of state and transport (mirgecom/gas_model.py:82-97,
 * mirgecom/eos.py, mirgecom/transport.py). */
typedef struct mg_gas {
  int32_t eos_kind;        /* 0 IdealSingleGas, 1 uiuc_7sp mixture          */
  int32_t transport_kind;  /* 0 none, 1 SimpleTransport, 2 PowerLawTransport,
                              3 MixtureAveragedTransport (uiuc_7sp mixture)  */
  int32_t nspecies;
  int32_t has_lewis;       /* power law / mixture-averaged: species_diffusivity
                              holds Lewis numbers                            */
  double gamma, gas_const; /* IdealSingleGas (eos.py:329)                   */
  double mu, mu_b, kappa;  /* SimpleTransport (transport.py:143-202)        */
  double pl_alpha, pl_beta, pl_sigma, pl_n; /* PowerLaw (transport.py:205-321)*/
  double species_diffusivity[8];
  /* MixtureAveragedTransport (transport.py:324-502): bulk-viscosity ratio
   * alpha, scaling factor, single-species threshold epsilon and the
   * diffusivity used there; needs mg_load_transport.                       */
  double ma_alpha, ma_factor, ma_epsilon, ma_singular_diffusivity;
} mg_gas;

/* Boundary data of Mengaldo-type boundaries (AdiabaticSlipBoundary,
 * IsothermalWallBoundary, AdiabaticNoslipWallBoundary, FarfieldBoundary,
 * PressureOutflowBoundary, ...; mirgecom/boundary.py:391-1857).  On boundary
 * nodes whose flags hold MG_BC_MENGALDO the gradients take CV_bc and T_bc as
 * boundary values (boundary.py:713-793) and the viscous boundary flux is
 * F_V(state_bc, grad_cv_bc, grad_t_bc) . n (boundary.py:630-710), with the
 * interior gradients conditioned by the other bits (slip: boundary.py:
 * 159-173; impermeable: 193-213; adiabatic: 186-190).  The inviscid flux
 * keeps using q_bnd/t_bnd (state_plus).  A NULL mg_wall (or NULL flags)
 * means no such nodes. */
#define MG_BC_MENGALDO 1
#define MG_BC_SLIP 2
#define MG_BC_IMPERMEABLE 4
#define MG_BC_ADIABATIC 8
typedef struct mg_wall {
  const double* q_bc;        /* [F][nbnd_nodes] CV_bc                       */
  const double* t_bc;        /* [nbnd_nodes] T_bc, the gradient value of T   */
  const double* t_state_bc;  /* [nbnd_nodes] temperature of state_bc (mixtures) */
  const int32_t* flags;      /* [nbnd_nodes] MG_BC_* bits                    */
  const int32_t* vol_index;  /* [nbnd_nodes] volume node behind each boundary node */
  const double* normals;     /* [dim][nbnd_nodes] outward unit normals        */
  double* work;              /* [(F+1)*dim][nbnd_nodes] scratch (caller-owned) */
} mg_wall;

/* Numerical inviscid flux kinds: inviscid_facial_flux_rusanov,
 * inviscid_facial_flux_hll, inviscid_facial_flux_central
 * (mirgecom/inviscid.py:105-287). */
#define MG_FLUX_RUSANOV 0
#define MG_FLUX_HLL 1
#define MG_FLUX_CENTRAL 2

/* Euler RHS, -div F_I with the numerical flux on faces
 * (replaces euler_operator, mirgecom/euler.py:313-416). temperature is
 * read for mixtures only (may be NULL otherwise); q_bnd/t_bnd hold the
 * exterior state of every boundary node (t_bnd may be NULL for ideal gas). */
mg_status mg_euler_rhs(const mg_mesh* mesh, const mg_gas* gas,
                       const double* q, const double* temperature,
                       const double* q_bnd, const double* t_bnd,
                       int32_t flux_kind, double* rhs, void* stream);

/* Mixed-precision mg_euler_rhs: the same right-hand side with the nodal
 * fluxes held in LDS relative to each element's first node and contracted
 * in FP32 (the weak divergence of that constant cancels its lift exactly),
 * physics, face fluxes, state and RHS in FP64. */
mg_status mg_euler_rhs_mixed(const mg_mesh* mesh, const mg_gas* gas,
                             const double* q, const double* temperature,
                             const double* q_bnd, const double* t_bnd,
                             int32_t flux_kind, double* rhs, void* stream);

/* DG gradients of the conserved fields and of T with the central flux on
 * interior faces and the exterior value (CV_bc / T_bc on Mengaldo nodes) on
 * boundary faces (replaces grad_cv_operator + grad_t_operator,
 * mirgecom/navierstokes.py:129-347). */
mg_status mg_ns_gradient(const mg_mesh* mesh, const mg_gas* gas,
                         const double* q, const double* temperature,
                         const double* q_bnd, const double* t_bnd,
                         const mg_wall* wall, double* grad, void* stream);

/* Navier-Stokes RHS div(F_V - F_I) from q and its gradients (replaces the
 * divergence part of ns_operator, mirgecom/navierstokes.py:533-575). */
mg_status mg_ns_rhs(const mg_mesh* mesh, const mg_gas* gas,
                    const double* q, const double* temperature,
                    const double* grad, const double* q_bnd,
                    const double* t_bnd, const mg_wall* wall, int32_t flux_kind,
                    int32_t inviscid_on, double* rhs, void* stream);

/* Fused first half of the Navier-Stokes RHS: the DG gradients of the
 * conserved fields and of T (as mg_ns_gradient) feed, in the same kernel,
 * the pointwise net flux F_I - F_V at every node, written to fnet
 * [F*dim][N] (F_I omitted if inviscid_on == 0).  The gradients themselves
 * are written only if grad != NULL; for the uiuc_7sp mixture grad
 * [(F+1)*dim][N] is required: the mixture path stores the gradients there
 * and forms the net fluxes in a node-parallel pass.  For boundary face nodes fnet_bnd
 * [F][nbnd_nodes] receives (F_I - F_V)(q_ext, grad q-) . n, the exterior side
 * of the boundary fluxes (mirgecom/boundary.py:944-970); on Mengaldo nodes
 * its viscous part is 2 F_V(bc).n - F_V(q-, grad q-).n, so that the central
 * average is F_V(bc).n.  Replaces
 * grad_cv_operator, grad_t_operator and the volume viscous_flux of
 * ns_operator (mirgecom/navierstokes.py:494-541). */
mg_status mg_ns_flux(const mg_mesh* mesh, const mg_gas* gas, const double* q,
                     const double* temperature, const double* q_bnd,
                     const double* t_bnd, const mg_wall* wall, int32_t inviscid_on,
                     double* fnet, double* fnet_bnd, double* grad, void* stream);

/* Second half: divergence of the nodal net fluxes with the numerical face
 * flux (Rusanov/central applied to the net normal fluxes, HLL to the
 * inviscid part plus the central viscous part), lift and inverse mass.
 * Replaces viscous_flux_on_element_boundary, inviscid_flux_on_element_boundary
 * and div_operator in ns_operator (mirgecom/navierstokes.py:544-563). */
mg_status mg_ns_div(const mg_mesh* mesh, const mg_gas* gas, const double* q,
                    const double* temperature, const double* fnet,
                    const double* q_bnd, const double* t_bnd,
                    const double* fnet_bnd, int32_t flux_kind,
                    int32_t inviscid_on, double* rhs, void* stream);

/* Mixed-precision variants of mg_ns_flux / mg_ns_div (the FP32 path of
 * the same right-hand side): gradients contracted in FP32 from traces held
 * relative to each element's first node, physics in FP64, nodal net flux
 * stored in FP32 with the pressure of node 0 (p_ref) removed from the
 * momentum-flux diagonal, divergence contracted in FP32 from fluxes held
 * relative to each element's first node, face fluxes formed in FP64.  Both
 * shifts leave the DG divergence unchanged in exact arithmetic.  The state q
 * and the RHS stay FP64.  fnet_bnd carries the p_ref shift as well.
 * Mixtures take two passes: FP32 gradients [(F+1)][dim][N] into grad
 * (caller-allocated staging, required for mixtures, unused otherwise), then
 * the net flux node by node. */
mg_status mg_ns_flux_mixed(const mg_mesh* mesh, const mg_gas* gas,
                           const double* q, const double* temperature,
                           const double* q_bnd, const double* t_bnd,
                           const mg_wall* wall, int32_t inviscid_on, float* fnet,
                           double* fnet_bnd, float* grad, void* stream);
mg_status mg_ns_div_mixed(const mg_mesh* mesh, const mg_gas* gas,
                          const double* q, const double* temperature,
                          const float* fnet, const double* q_bnd,
                          const double* t_bnd, const double* fnet_bnd,
                          int32_t flux_kind, int32_t inviscid_on, double* rhs,
                          void* stream);

/* Mixture temperature by a fixed number of Newton iterations from a seed
 * (replaces PyrometheusMixture.temperature, mirgecom/eos.py:777-803, and
 * get_temperature, mirgecom/thermochemistry.py:126-167). q is [F][n]. */
mg_status mg_mixture_temperature(const mg_gas* gas, int32_t dim, int64_t n,
                                 const double* q, const double* t_seed,
                                 int32_t niter, double* t_out, void* stream);

/* The same Newton iteration run to a tolerance instead of a fixed count:
 * per node until an update |dT| <= tol (K) or max_iter updates; iters
 * (optional, [n]) receives the updates taken, -1 where the last one still
 * exceeded tol.  No counterpart in the reference, which iterates a fixed
 * number of times (mirgecom/thermochemistry.py:149-167). */
mg_status mg_mixture_temperature_tol(const mg_gas* gas, int32_t dim, int64_t n,
                                     const double* q, const double* t_seed, double tol,
                                     int32_t max_iter, double* t_out, int32_t* iters,
                                     void* stream);

/* Species mass sources W_k * omega_k into src [nspecies][n]
 * (replaces get_species_source_terms, mirgecom/eos.py:858-875). */
mg_status mg_species_source(const mg_gas* gas, int32_t dim, int64_t n,
                            const double* q, const double* temperature,
                            double* src, void* stream);

/* The same sources added into the species rows of a right-hand side
 * rhs [F][n] (rows 2 + dim ..), in place: the reacting RHS
 * euler_operator(...) + get_species_source_terms(...) without a separate
 * source array or add pass. */
mg_status mg_species_source_add(const mg_gas* gas, int32_t dim, int64_t n,
                                const double* q, const double* temperature,
                                double* rhs, void* stream);

/* Node-local stable timestep into dt [K*np] from the element characteristic
 * lengths lengthscale [K]: h/(|v|+c) (viscous == 0; replaces
 * get_inviscid_timestep, mirgecom/inviscid.py:379-422) or
 * h/(|v|+c+(mu/rho+max_k d_k)/h) (viscous != 0; replaces
 * get_viscous_timestep, mirgecom/viscous.py:547-607). */
mg_status mg_local_timestep(const mg_mesh* mesh, const mg_gas* gas, const double* q,
                            const double* temperature, const double* lengthscale,
                            int32_t viscous, double* dt, void* stream);

/* Classical RK4 stage helpers over n values (mirgecom/integrators/
 * explicit_rk.py:31-38): out = y + a*k ; acc = acc + w*k (acc may be NULL). */
mg_status mg_rk_axpy(int64_t n, const double* y, const double* k, double a,
                     double* out, double* acc, double w, void* stream);

/* Bound-preserving limiter (mirgecom/limiter.py:39-125,
 * bound_preserving_limiter): per element, u <- avg + theta (u - avg), theta the
 * largest factor in [0, 1] that keeps the element's nodal values inside
 * [mmin, mmax] (the upper bound only with MG_LIMIT_MAX).  field and out are
 * [nfields][nelements][np] (out may equal field); weights[np] are the row sums
 * of the reference mass matrix, which give the cell average on an affine
 * simplex.  MG_LIMIT_MODIFY_AVERAGE clips the average into the bounds first
 * (the reference's modify_average).  np is a simplex node count: 3, 4, 6, 10,
 * 15, 20, 21, 35 or 56. */
#define MG_LIMIT_MAX 1
#define MG_LIMIT_MODIFY_AVERAGE 2
mg_status mg_bound_limiter(int32_t np, int64_t nelements, int32_t nfields,
                           const double* weights, const double* field, double* out,
                           double mmin, double mmax, int32_t flags, void* stream);

/* Operators of the quadrature discretization (overintegration,
 * DISCR_TAG_QUAD; mirgecom/discretization.py:65-83): nq volume points and
 * nfq points per face, device arrays stored transposed (index of the
 * point or node the kernel's lanes run over last):
 *   interp_t[i][m]      = I[m][i],  I: nodal values -> volume points
 *   weak_diff_t[r][m][i] = (M^-1 D_q,r^T W)[i][m]
 *   face_interp_t[f][j][m] = face nodes (face_mask order) -> face points
 *   lift_t[f*nfq+m][i]  = (M^-1 E_q)[i][f*nfq+m] */
typedef struct mg_quad {
  int32_t nq, nfq;
  const double* interp;       /* [np][nq]            */
  const double* weak_diff;    /* [dim][nq][np]       */
  const double* face_interp;  /* [nfaces][nfp][nfq]  */
  const double* lift;         /* [nfaces*nfq][np]    */
} mg_quad;

/* Euler RHS on the quadrature discretization (euler_operator with
 * quadrature_tag=DISCR_TAG_QUAD, mirgecom/euler.py:367-416): the state is
 * interpolated to the volume points, where the inviscid flux is formed, and
 * both face traces to the face points, where the numerical flux (flux_kind)
 * is formed; weak divergence and lift integrate them with the quadrature
 * weights.  Ideal gas (passive species allowed) or the uiuc_7sp mixture:
 * for the mixture the temperature at every point is the niter-step Newton
 * inversion seeded with the interpolated nodal temperature [N]
 * (project_fluid_state, mirgecom/gas_model.py:494-588).  q_bnd_quad
 * [F][nbnd_faces*nfq] (and t_bnd_quad [nbnd_faces*nfq] for mixtures) hold the
 * boundary exterior states at the boundary faces' points (face by face, in
 * the order of the boundary nodes); NULL without boundaries. */
mg_status mg_euler_rhs_quad(const mg_mesh* mesh, const mg_gas* gas,
                            const mg_quad* quad, const double* q,
                            const double* temperature, const double* q_bnd_quad,
                            const double* t_bnd_quad, int32_t niter, int32_t flux_kind,
                            double* rhs, void* stream);

/* DG gradients of the conserved fields and of T on the quadrature
 * discretization (grad_cv_operator + grad_t_operator with
 * quadrature_tag=DISCR_TAG_QUAD, mirgecom/navierstokes.py:129-347): the
 * state interpolated to the volume and face points, T from the EOS there
 * (mixtures: niter Newton steps seeded with the interpolated nodal
 * temperature), the central value at interior face points and the exterior
 * state at boundary points (prescribed / dummy boundaries: q_bnd_quad,
 * t_bnd_quad as for mg_euler_rhs_quad), integrated with the quadrature
 * weights.  grad [F+1][dim][N] (T last), the layout of mg_ns_gradient.
 * The gas needs transport; nspecies 0 or 7 (the mixture: 7). */
mg_status mg_ns_gradient_quad(const mg_mesh* mesh, const mg_gas* gas,
                              const mg_quad* quad, const double* q,
                              const double* temperature, const double* q_bnd_quad,
                              const double* t_bnd_quad, int32_t niter, double* grad,
                              void* stream);

/* Navier-Stokes RHS div(F_V - F_I) on the quadrature discretization
 * (ns_operator with quadrature_tag=DISCR_TAG_QUAD,
 * mirgecom/navierstokes.py:350-575): the viscous and (inviscid_on != 0)
 * inviscid fluxes at the volume points from the point states and the nodal
 * gradients grad [F+1][dim][N] interpolated there (op.project); at the face
 * points the numerical inviscid flux (flux_kind) and the central viscous
 * flux of both sides' states and gradient traces (boundary points: the
 * exterior state with the interior gradients, viscous.py:431-544); weak
 * divergence and lift with the quadrature weights.  Arguments as for
 * mg_ns_gradient_quad. */
mg_status mg_ns_rhs_quad(const mg_mesh* mesh, const mg_gas* gas, const mg_quad* quad,
                         const double* q, const double* temperature,
                         const double* grad, const double* q_bnd_quad,
                         const double* t_bnd_quad, int32_t niter, int32_t flux_kind,
                         int32_t inviscid_on, double* rhs, void* stream);

/* Hybridized summation-by-parts operators of the entropy-stable path (on
 * the quadrature of an mg_quad; nh = nq + nfaces*nfq hybridized points:
 * volume points, then each face's points), device arrays:
 *   proj_t[m][i]   = (M^-1 V_q^T W)[i][m]          [nq][np]
 *   vh_t[i][h]     = V_h[h][i]                     [np][nh]
 *   skew_t[r][j][h] = (Q_r - Q_r^T)[h][j], Q_r = W V_q D_r P_q   [dim][nq][nq]
 *   ext[m][h]      = E[m][h], E = V_f P_q          [nfaces*nfq][nq]
 *   ext_col[h][m]  = E[m][h]                       [nq][nfaces*nfq]
 *   apply_t[h][i]  = (M^-1 V_h^T)[i][h]            [nh][np]
 *   face_weights   = the face rule's weights       [nfq]
 * and the same operators along the cyclic diagonals (optional, used when
 * nq = 64 and nfaces*nfq = 192, i.e. 3-D p = 3 at degree 6, where each
 * two-point flux is then evaluated once for both of its rows):
 *   skew_diag[r][s][h] = (Q_r - Q_r^T)[h][(h + s) mod nq]   [dim][nq][nq]
 *   ext_diag[s][m]     = E[m][(m + s) mod nq]               [nq][nfaces*nfq] */
typedef struct mg_esdg {
  int32_t nh, pad0;
  const double* proj;
  const double* vh;
  const double* skew;
  const double* ext;
  const double* ext_col;
  const double* apply;
  const double* face_weights;
  const double* skew_diag;
  const double* ext_diag;
} mg_esdg;

/* Entropy-stable Euler RHS by flux differencing (euler_operator with
 * use_esdg=True and quadrature_tag=DISCR_TAG_QUAD, mirgecom/euler.py:108-310)
 * for the ideal gas: entropy variables projected onto the element space
 * (stored in ev_work [F][N], caller-owned), the state at the hybridized
 * points from them, Chandrashekar's entropy-conserving two-point flux
 * (mirgecom/inviscid.py:451-518) in (Q_h - Q_h^T) o F_S, and the
 * entropy-stable face flux (inviscid.py:601-711): with dissipation != 0 the
 * Rusanov penalty alpha = max wavespeed, else none.  q_bnd_quad as for
 * mg_euler_rhs_quad. */
mg_status mg_euler_rhs_esdg(const mg_mesh* mesh, const mg_gas* gas,
                            const mg_quad* quad, const mg_esdg* esdg, const double* q,
                            const double* q_bnd_quad, int32_t dissipation,
                            double* ev_work, double* rhs, void* stream);

/* Entropy-stable Euler RHS for the uiuc_7sp mixture (euler_operator with
 * use_esdg=True and quadrature_tag=DISCR_TAG_QUAD, mirgecom/euler.py:108-310,
 * with Renac's entropy-conserving two-point flux, mirgecom/inviscid.py:
 * 521-598).  Volume points: the interpolated state and its temperature by
 * niter Newton steps from the interpolated nodal temperature [N]; with
 * es_volume != 0 taken through its entropy variables (with the state's own
 * p and gamma) and back and made again (project_fluid_state(...,
 * entropy_stable=True), gas_model.py:494-588; euler_operator's path), else
 * as it is (the states ns_operator hands to its inviscid operator).  Its
 * entropy variables are L2-projected into ev_work [F][N].  Hybridized points:
 * conserved variables from the projected variables and the interpolated
 * nodal gamma(cv, T) (gas_model.py:1000-1023), T by Newton from the
 * interpolated nodal T; interior faces: the neighbour's projected variables
 * with this element's gamma and the neighbour's T as the seed; boundary
 * faces: the entropy-stable projected interior state and q_bnd_quad with its
 * temperatures t_bnd_quad [nbnd_faces*nfq].  Face flux: Renac's flux . n,
 * with dissipation != 0 minus alpha/2 (q+ - q-), alpha the larger |u| + c.
 * 2-D p = 1..4, 3-D p = 1..3. */
mg_status mg_euler_rhs_esdg_mixture(const mg_mesh* mesh, const mg_gas* gas,
                                    const mg_quad* quad, const mg_esdg* esdg,
                                    const double* q, const double* temperature,
                                    const double* q_bnd_quad, const double* t_bnd_quad,
                                    int32_t niter, int32_t dissipation, int32_t es_volume,
                                    double* ev_work, double* rhs, void* stream);

/* Allocation-free variant of mg_euler_rhs_esdg.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The projected entropy variables that mg_euler_rhs_esdg keeps in the
 * caller's ev_work live in a file-scope device array of the library instead:
 * 2^23 doubles (64 MiB) per GPU, held until the process ends, so
 * (2 + dim + nspecies) * nelements * np must not exceed 8,388,608 (the
 * 24^3 x 6-tet p = 3 ideal-gas box fits); larger states are rejected with
 * MG_ERR_ARG.  The call allocates nothing and frees nothing.  All calls on
 * one GPU share that array: calls that overlap on one GPU (other streams or
 * host threads) race and must be serialised by the caller.  Results equal
 * mg_euler_rhs_esdg's bit for bit. */
mg_status mg_euler_rhs_esdg_noalloc(const mg_mesh* mesh, const mg_gas* gas,
                                    const mg_quad* quad, const mg_esdg* esdg,
                                    const double* q, const double* q_bnd_quad,
                                    int32_t dissipation, double* rhs, void* stream);

/* Upload the uiuc_7sp mechanism constants (NASA7 tables, molecular weights,
 * gas constant, standard pressure, SI Arrhenius parameters; 130 doubles in
 * the order of mirgecom_amd.mechanisms.packed_mechanism) to the device.
 * Needed once per process before any mixture call; it replaces the
 * pyrometheus code generation of the mechanism class
 * (mirgecom/thermochemistry.py:186-219). */
mg_status mg_load_mechanism(const double* packed, int64_t count);

/* Upload the pure-species transport fits of the uiuc_7sp mixture (308
 * doubles in the order of mirgecom_amd.transport_fits.packed_transport:
 * sqrt(mu_k/sqrt(T)), lambda_k/sqrt(T), p D_kj/T^1.5 as degree-4 polynomials
 * in ln T, and the Wilke factors).  Needed once per process before any call
 * with transport_kind 3; it replaces the transport part of the pyrometheus
 * mechanism class that MixtureAveragedTransport evaluates
 * (mirgecom/transport.py:391-502). */
mg_status mg_load_transport(const double* packed, int64_t count);

/* Transport variables at n nodes (mirgecom/transport.py:66-83,
 * TransportModel.transport_vars): mu, mu_B, kappa [n] and the species
 * diffusivities [nspecies][n] from the conserved state q [F][n] (and the
 * temperature [n] for mixtures), evaluated exactly as inside the
 * Navier-Stokes kernels.  Ideal gas without species, or the uiuc_7sp mixture. */
mg_status mg_transport_vars(const mg_gas* gas, int32_t dim, int64_t n,
                            const double* q, const double* temperature, double* mu,
                            double* mu_b, double* kappa, double* diff, void* stream);

/* {{{ Learned chemistry surrogate (no counterpart in the reference: it
 * stands in for get_species_source_terms, mirgecom/eos.py:858-875, on the
 * uiuc_7sp mixture).
 *
 * Network 10 -> 256 -> 256 -> 256 -> 7 with GELU; inputs (log T, log p,
 * Y_1..Y_7, dt) normalised, outputs the normalised change dY_k over a
 * chemistry substep dt.  BF16 weights and activations on the matrix cores,
 * FP32 accumulation.  params: FP32 master array of MG_MLP_NPARAM values
 * (W_l as [K_in][N_out] row-major then b_l; widths padded to 32 -> 256 ->
 * 256 -> 256 -> 32); wbf16: the BF16 copies of the weights, MG_MLP_NBF16
 * values, kept in step by mg_mlp_adam / mg_mlp_sync_bf16. */
#define MG_MLP_NPARAM 148256
#define MG_MLP_NBF16 294912

typedef struct mg_chem_norm {
  double in_mean[10], in_scale[10];   /* x = (raw - mean) / scale            */
  double out_mean[7], out_scale[7];   /* dY_k = out_k * scale_k + mean_k     */
  double lo[10], hi[10];              /* gate: raw features inside [lo, hi]   */
  double check_tol;                   /* gate: sampled |S - S_exact| bound in
                                         output scales (times rho / dt)       */
  int32_t sample_period, sample_offset; /* gate: nodes g with g % period ==
                                         offset get the exact-rate check (0:
                                         no sampled check)                    */
} mg_chem_norm;

/* Constant-pressure adiabatic reactors (the 0-D reactor of
 * test/test_chemistry.py:51-110, integrated with classical RK4 at step dt, T
 * from the conserved enthalpy by Newton after every stage): n reactors from
 * t0 [n], p [n], y0 [7][n].  t_ign [n] receives the first time T exceeds
 * t0 + t_rise (-1 if not within nmax steps).  With nsamp > 0, states
 * [nsamp][n][10] = (T, p, Y_1..7, dt) are recorded at equal step spacing
 * over min(2 t_ign, nmax dt) (nmax dt without ignition), and dy [nsamp][n][7]
 * the change of Y over the following step. */
mg_status mg_reactor_trajectories(int64_t n, const double* t0, const double* p,
                                  const double* y0, double dt, int64_t nmax,
                                  double t_rise, int32_t nsamp, double* t_ign,
                                  double* states, double* dy, void* stream);

/* Change of Y over dt (nsub RK4 steps at constant pressure and enthalpy)
 * from states [n][10] = (T, p, Y_1..7, ...): dy [n][7]. */
mg_status mg_chem_substep(int64_t n, const double* states, double dt, int32_t nsub,
                          double* dy, void* stream);

/* Network inputs from states [n][10]: x [n][32] BF16 (raw uint16 bits);
 * with dy [n][7] given, the normalised targets [n][32] FP32. */
mg_status mg_chem_features(int64_t n, const double* states, const double* dy,
                           const mg_chem_norm* norm, uint16_t* x, float* target,
                           void* stream);

/* Training.  batch: a positive multiple of 256.  work: a caller-allocated
 * device workspace of mg_mlp_workspace_bytes(batch) bytes that carries the
 * activations from forward to backward and the gradient from backward to
 * adam.  forward writes the network outputs (normalised, [batch][32] FP32)
 * to out if out != NULL; backward takes the loss L = sum_{k<7} (out - target)^2
 * / (7 batch), writes its per-block partial sums to loss_part
 * ((32 batch + 255) / 256 floats; may be NULL) and forms dL/dparams; adam
 * (Kingma & Ba, bias-corrected, step from 1) updates params, m, v and the
 * BF16 copies, and copies the gradient to grad_out if != NULL. */
int64_t mg_mlp_param_count(void);
int64_t mg_mlp_bf16_count(void);
int64_t mg_mlp_workspace_bytes(int64_t batch);
mg_status mg_mlp_forward(int64_t batch, const uint16_t* x, const float* params,
                         const uint16_t* wbf16, void* work, float* out, void* stream);
mg_status mg_mlp_backward(int64_t batch, const uint16_t* x, const float* target,
                          const float* params, const uint16_t* wbf16, void* work,
                          float* loss_part, void* stream);
mg_status mg_mlp_adam(float* params, float* m, float* v, const void* work, int64_t batch,
                      int32_t step, float lr, float beta1, float beta2, float eps,
                      uint16_t* wbf16, float* grad_out, void* stream);
/* Re-emit the BF16 copies from params (after the host set params). */
mg_status mg_mlp_sync_bf16(const float* params, uint16_t* wbf16, void* stream);

/* Species sources from the surrogate for n nodes (n a multiple of np, the
 * nodes of whole elements in order): S_k = rho dY_k / dt.  Gate, per
 * element: a node whose raw features leave [lo, hi], or a sampled node whose
 * S misses the exact rates by more than check_tol output scales, sends the
 * whole element to the exact rates (mg_species_source's).  add != 0: out is
 * a right-hand side [F][n] whose species rows receive += S; else out is
 * [7][n].  fallback_count (optional, device int) is incremented by the
 * number of elements that took the exact rates. */
mg_status mg_chem_surrogate_source(const mg_gas* gas, int32_t dim, int32_t np, int64_t n,
                                   const double* q, const double* temperature, double dt,
                                   const float* params, const uint16_t* wbf16,
                                   const mg_chem_norm* norm, int32_t add, double* out,
                                   int32_t* fallback_count, void* stream);
/* }}} */

/* {{{ Batched ensemble of independent cases (no counterpart in the
 * reference, which runs one driver per case: mirgecom/steppers.py:350-449).
 *
 * Every case is a periodic 3-D uiuc_7sp mixture with a non-mixture-averaged
 * transport (its own mesh, order 1..4 and gas parameters, e.g. a Reynolds
 * number through mu), evaluated with the mixed-precision Navier-Stokes path
 * (mg_ns_flux_mixed / mg_ns_div_mixed with the Rusanov flux) after the
 * Newton temperature from t_seed; then the species sources are added into
 * rhs: the learned surrogate (mg_chem_surrogate_source) when mlp_params is
 * given, else the exact rates.  Results equal those of the single-case entry
 * points bit for bit. */
typedef struct mg_ens_case {
  mg_mesh mesh;            /* prepared (mg_mesh_prepare), nbnd_nodes == 0   */
  mg_gas gas;
  const double* q;         /* [F][N] state to evaluate                       */
  const double* t_seed;    /* [N] Newton seed                                */
  double* temperature;     /* [N] Newton result                              */
  float* fnet;             /* [F*3][N] FP32 nodal net flux (work)           */
  float* grad;             /* [(F+1)*3][N] FP32 gradients (work)            */
  double* rhs;             /* [F][N] right-hand side (output)                */
} mg_ens_case;

/* fused != 0: one launch per phase for all orders (the order is selected per
 * work item); else one per order.  persistent != 0: workgroups stride over
 * the work items, wg_per_cu[p] of them per compute unit for order p's
 * launches ([0]: the fused launch); else one workgroup per item. */
typedef struct mg_ens_tuning {
  int32_t fused;
  int32_t persistent;
  int32_t wg_per_cu[5];
} mg_ens_tuning;

/* Host handle of a prepared ensemble (filled by mg_ensemble_prepare; the
 * tables and item lists live in the caller's device buffer dev). */
typedef struct mg_ens_plan {
  void* dev;
  int32_t ncases, fused, niter, cus, surrogate, persistent;
  int64_t grad_begin[5], grad_count[5], div_begin[5], div_count[5];
  int32_t grad_seg_begin[5], grad_seg_count[5], div_seg_begin[5], div_seg_count[5];
  int64_t pt_count, src_count;
  int32_t grid_grad[5], grid_div[5], grid_pt, grid_src;
  int64_t off_grad_tab, off_div_tab, off_pt_tab, off_src_tab;
  int64_t off_grad_items, off_div_items, off_pt_items, off_src_items;
} mg_ens_plan;

/* Device bytes of the plan of these cases (-1 on invalid cases). */
int64_t mg_ensemble_plan_bytes(int32_t ncases, const mg_ens_case* cases);

/* Build the per-case descriptors and work items into dev (synchronises the
 * stream once; not capturable).  niter: Newton iterations; dt: the chemistry
 * substep of the surrogate; fallback_count (optional device int) counts the
 * elements the surrogate's gate sends to the exact rates; tuning may be
 * NULL (fused, one workgroup per item). */
mg_status mg_ensemble_prepare(int32_t ncases, const mg_ens_case* cases, int32_t niter,
                              double dt, const float* mlp_params,
                              const uint16_t* mlp_wbf16, const mg_chem_norm* norm,
                              int32_t* fallback_count, const mg_ens_tuning* tuning,
                              void* dev, int64_t dev_bytes, mg_ens_plan* plan,
                              void* stream);

/* One right-hand side of every case of the plan: four launches (Newton,
 * gradients + net flux, divergence, sources) with the fused tuning, 2 + 2x4
 * otherwise.  Capturable in a HIP graph. */
mg_status mg_ensemble_ns_rhs(const mg_ens_plan* plan, void* stream);
/* }}} */

/* Message of the last failed call on this thread. */
const char* mg_last_error(void);

/* Library ABI version (4: mg_esdg carries skew_diag and ext_diag). */
int32_t mg_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MGDG_H */
