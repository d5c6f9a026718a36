// Cut-cell geometry of the pressure grid on the GPU (level set -> edge fractions -> face and
// cell types -> face fractions -> cut volumes -> operator coefficients).
#pragma once

#include <hip/hip_runtime.h>

namespace lsx {
namespace geo {

// Reference cell type codes (src/Base/LS_STAG_Mesh/cut_cells.h:26-168): interior cut cells
// carry an 8-digit decimal code of their fluid corners, domain border cells negative codes.
enum CellCode : int {
  SolidCell = 0,
  FluidCell = 11111111,
  T_Border = -2, B_Border = -3, N_Border = -4, S_Border = -5, E_Border = -6, W_Border = -7,
  TN_Edge = -8, TS_Edge = -9, TE_Edge = -10, TW_Edge = -11, BN_Edge = -12, BS_Edge = -13,
  BE_Edge = -14, BW_Edge = -15, NE_Edge = -16, NW_Edge = -17, SE_Edge = -18, SW_Edge = -19,
  TNE_Corner = -20, TNW_Corner = -21, TSE_Corner = -22, TSW_Corner = -23,
  BNE_Corner = -24, BNW_Corner = -25, BSE_Corner = -26, BSW_Corner = -27
};

enum BodyKind : int { BODY_SPHERE = 0, BODY_ELLIPSOID = 1, BODY_BEZIER6 = 2 };

struct Body {
  int kind;
  double c[3];                   // centre
  double a[3];                   // radius (sphere: a[0]) or semi-axes (Bezier: start of the search)
  const double* ctrl = nullptr;  // Bezier body: 288 control values in device memory
};

// cut cells may appear for x in [hs, hf), y in [vs, vf), every z plane
struct Box {
  int hs, hf, vs, vf;
};

int rule_count();
void level_set(const double* cx, const double* cy, const double* cz, int N, int M, int K, Box box, Body body,
               double* phi, hipStream_t s);
// phi = phi0 after the oscillation fix (tmp: n doubles of scratch, changed: one device int);
// synchronises the stream; returns the number of passes (-1 on a HIP error)
int oscillation_fix(const double* phi0, double* phi, double* tmp, unsigned char* flag, int* changed, int N, int M,
                    int K, Box box, hipStream_t s);
void alpha(const double* phi, int N, int M, int K, Box box, double* ax, double* ay, double* az, hipStream_t s);
void types(const double* phi, int N, int M, int K, Box box, int* type, int* sx, int* sy, int* sz, hipStream_t s);
void theta(const double* ax, const double* ay, const double* az, const int* sx, const int* sy, const int* sz, int N,
           int M, int K, Box box, double* tx, double* ty, double* tz, hipStream_t s);
void vof(const int* type, const double* ax, const double* ay, const double* az, const double* tx, const double* ty,
         const double* tz, const double* dx, const double* dy, const double* dz, int N, int M, int K, double* v,
         hipStream_t s);
// momentum operator of velocity component dir (0 u, 1 v, 2 w) on its staggered mesh
// ((N-1) x M x K, N x (M-1) x K or N x M x (K-1)) as 7 planes [b,s,w,p,e,n,t] of that mesh's size;
// a box without cut cells (the caller checks)
void helmholtz7(const int* type, const double* tx, const double* ty, const double* tz, const double* v,
                const double* dx, const double* dy, const double* dz, int N, int M, int K, int dir, double nu,
                double coef_time, int flow, double* coef, hipStream_t s);
void laplace7(const int* type, const double* tx, const double* ty, const double* tz, const double* v, const double* dx,
              const double* dy, const double* dz, int N, int M, int K, int pin, double* coef, hipStream_t s);

// boundary velocities of the pressure equation's flux (flow_task_description.h:79-96)
struct FlowBc {
  double U_E, V_S, V_N, W_B, W_T;
  int pin;  // b = 0 at cells N-1 and N*M-1
};
// u = (G dp) inv_M + u*, v and w likewise (velocity part of the projection correction)
void correct(const int* type, const double* tx, const double* ty, const double* tz, const double* vof, const double* dx,
             const double* dy, const double* dz, int N, int M, int K, const double* dp, const double* us,
             const double* vs, const double* ws, double* u, double* v, double* w, hipStream_t s);
// p = p + dp * coef
void pressure_update(double* p, const double* dp, double coef, long n, hipStream_t s);
// b = D u* + flux on the base grid (u_west: one inflow velocity per x-row; u_out: x-face field
// whose last interior face feeds an outflow east boundary, or null for the fixed U_E)
void pressure_rhs(const int* type, const double* tx, const double* ty, const double* tz, const double* dx,
                  const double* dy, const double* dz, int N, int M, int K, const double* u, const double* v,
                  const double* w, const double* u_west, const double* u_out, FlowBc bc, double* b, hipStream_t s);

// ---- objective and shape gradient ----------------------------------------------------------
enum Objective : int { OBJ_NONE = 0, OBJ_DRAG_X = 1 };

// read-only view of one geometry's fields
struct GeoView {
  const int* type;
  const double *tx, *ty, *tz, *vof;
};
// the flow fields the pressure right-hand side is assembled from (device pointers)
struct RhsArgs {
  const double *u, *v, *w, *u_west, *u_out;
  FlowBc bc;
};

// J = sum over cut cells of (theta_x[i-1] - theta_x[i]) dy dz p[i]; dJdp (may be null) gets the
// coefficients; partial: 256 doubles of device scratch.  Synchronises the stream.
double drag(const int* type, const double* tx, const double* dy, const double* dz, int N, int M, int K, Box box,
            const double* p, double* dJdp, double* partial, hipStream_t s);
// The same J and dJdp with no allocation and no host synchronisation: the partial sums go to
// `partial` (the address drag_scratch returns for the current device), J (device pointer)
// receives the sum in the order drag() adds them.
void drag_noalloc(const int* type, const double* tx, const double* dy, const double* dz, int N, int M, int K, Box box,
                  const double* p, double* dJdp, double* partial, double* J, hipStream_t s);
// Address of the library's drag scratch array on the current device (looked up once per device
// with hipGetSymbolAddress); 0, or a negative value when the lookup fails.
int drag_scratch(double** out);
// phi_p = phi_raw + h, phi_m = phi_raw - h at the in-box corners of class cls (x%6 + 6 (y%6) + 36 (z%6))
void perturb_class(const double* phi_raw, int N, int M, int K, Box box, int cls, double h, double* phi_p,
                   double* phi_m, hipStream_t s);
// dF[i] = (F_i(+) - F_i(-)) / 2h with F_i = J_i + lambda_i (b_i - (A p)_i); rhs null: b fixed
void lagrangian_diff(const GeoView& gp, const GeoView& gm, const double* dx, const double* dy, const double* dz, int N,
                     int M, int K, Box box, int pin, const double* p, const double* lam, int objective,
                     const RhsArgs* rhs, double h, double* dF, hipStream_t s);
// g[c] = sum of dF over cells c + [-2, 3]^3 for the corners c of class cls
void corner_gather(const double* dF, int N, int M, int K, Box box, int cls, double* g, hipStream_t s);
// out[288] = sum over corners of g dphi/dctrl through the closest point (hf, huv, hw: n, 2n, 3n scratch)
void bezier_chain(const double* g, const double* cx, const double* cy, const double* cz, int N, int M, int K, Box box,
                  Body body, int* hf, double* huv, double* hw, double* out, hipStream_t s);

}  // namespace geo
}  // namespace lsx
