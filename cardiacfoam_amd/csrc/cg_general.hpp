// The monodomain diffusion step on an arbitrary polyhedral mesh (general_mesh.hip): the
// mesh as OpenFOAM hands it over (lduAddressing owner/neighbour, face area vectors and
// centres, cell centres and volumes, boundary faces), fvc::laplacian / fvm::laplacian with
// a tensor conductivity and Gauss linear corrected (non-orthogonal correction included),
// zeroGradient on every boundary face.
#pragma once
#include <string>
#include <hip/hip_runtime.h>

#include "cg_internal.hpp"

namespace cg {

struct GeneralMesh;

// owner/nbr [nf] (lduAddressing order, owner < nbr), Sf/Cf [nf][3], C [n][3], vol [n],
// boundary faces b_owner [nb], b_Sf [nb][3] (all host).  Returns 0, 1 on bad input (err
// says why), 5 when device memory runs out.
int general_create(GeneralMesh** out, int64_t n, int64_t nf, const int32_t* owner,
                   const int32_t* nbr, const double* Sf, const double* Cf, const double* C,
                   const double* vol, int64_t nb, const int32_t* b_owner, const double* b_Sf,
                   std::string* err);
void general_destroy(GeneralMesh* G);
// face coefficients from the per-cell tensors sigma9 [n][9] (host, row-major)
int general_set_sigma(GeneralMesh* G, const double* sigma9, std::string* err);
// suggestExplicitDeltaT (myocardiumDomain.C:532-550) on the internal faces
double general_explicit_dt(const GeneralMesh* G, double chiCm, double max_co);
const int32_t* general_owner(const GeneralMesh* G);
const int32_t* general_neighbour(const GeneralMesh* G);
int64_t general_faces(const GeneralMesh* G);

// constants of one step (ddt coefficients c V^{n+1} = c0 V^n - c00 V^{n-1} + ...)
struct GeneralStep {
    double dt, chiCm, c, c0, c00, t_old;
    double act_v;       // activation threshold (V), 0 the reference's
    ActiveBoxes boxes;
    const double* src;  // source field [n] (device, A/m^3) in place of the boxes, or null
};

// the general-mesh description of a uniform hex block (blockMesh order, walls as boundary
// faces), for the diffusion-only entry on a hex tissue
int general_from_hex(GeneralMesh** out, const HexGeom& g, std::string* err);

// explicit step: Vnew = (c0 Vn - c00 Vm1 + dt/chiCm (lap(Vn) - chiCm Iion + src)) / c and the
// activation update.  Asynchronous on s.
hipError_t general_explicit(GeneralMesh* G, const double* Vn, const double* Vm1,
                            double* Vnew, const double* Iion, uint8_t* flag, double* tact,
                            const GeneralStep& st, hipStream_t s);

// fvMatrix of the implicit step, volume-integrated as OpenFOAM assembles it:
// diag = chiCm c vol/dt + sum a_f, upper = -a_f (host copies for the lduMatrix solver)
void general_matrix(const GeneralMesh* G, double chiCm, double c, double dt, double* diag,
                    double* upper);
// b0 = vol (chiCm/dt (c0 Vn - c00 Vm1) - chiCm Iion + src); asynchronous on s
hipError_t general_rhs0(GeneralMesh* G, const double* Vn, const double* Vm1,
                        const double* Iion, double* b0, const GeneralStep& st, hipStream_t s);
// b = b0 + explicit correction flux sum of psi (the non-orthogonal and tensor parts of the
// face fluxes); asynchronous on s
hipError_t general_rhs_corr(GeneralMesh* G, const double* psi, const double* b0, double* b,
                            hipStream_t s);
// initial guess of the solve: Vn, or 2 Vn - Vm1 (extrapolate); asynchronous on s
hipError_t general_guess(GeneralMesh* G, const double* Vn, const double* Vm1, double* psi,
                         bool extrapolate, hipStream_t s);
// activation update after an implicit solve (myocardiumDomain.C:345-374)
hipError_t general_activation(GeneralMesh* G, const double* Vn, const double* Vnew,
                              uint8_t* flag, double* tact, double t_old, double dt,
                              double vthr, hipStream_t s);
// pseudo-ECG (pseudoECGSolver.C:36-83) of V at n_el electrodes pos [E][3] (device): the
// conductivity per cell sig ([6][n] when sym, else [9][n]; device) or the uniform s9 (host);
// partial [E][nblocks] scratch, out [E] (device).  Asynchronous on s.
hipError_t general_pseudo_ecg(GeneralMesh* G, const double* V, const double* sig, int sym,
                              const double* s9, int n_el, const double* pos, int nblocks,
                              double* partial, double* out, hipStream_t s);
// div(sigma grad V) per unit volume into lap (device), for tests and adapters
hipError_t general_laplacian(GeneralMesh* G, const double* V, double* lap, hipStream_t s);

} // namespace cg
