// mf_plugin_api.h -- the case plugin API in the reference's signature style.
//
// A mesoflow case supplies its physics as compile-time functions selected by
// its Make.package (tests/Cartesian/ShockTube_N2/Make.package:1-2): thermo
// (thermo.H), transport and sources (userfuncs.H:30-54, transport.H:12-21,
// chemistry.H:13-22), the ext_dir boundary state (externbc.H:13-44) and the
// species tables (species.cpp). A case for this library writes the same
// functions, with the same names, argument meaning and component indices,
// in one header `<case>/mflo_case.H` that defines `namespace <case> { struct
// Case { ... }; }` (the directory name is the namespace). The build compiles
// every listed case into the library as one more plugin family
// (`make -C mesoflow_amd/csrc USER_PLUGINS="dir1 dir2"`; the cases under
// mesoflow_amd/plugins are built by default), with no change to the library's
// sources. Host and device code: the header compiles with hipcc and with a
// plain C++ compiler.
//
// Differences from the reference's form, all forced by the library keeping
// no managed globals:
//  * every function takes the context's parameter table (`const Params &`,
//    the C-ABI's mf_plugin_params) as its last argument -- the reference's
//    ParmParse-filled managed globals (e.g. mflo_user_funcs::fs_vel) become
//    fields of that table (in_vel, ...), its species tables are filled by
//    Case::init_species at context creation;
//  * `phi` in compute_fluid_transport / compute_spec_dcoeff /
//    compute_fluid_source / compute_spec_source is a view of the one cell
//    (i, j, k) the call is about (components 0 .. TOTAL_NVARS-1; every plugin
//    of the reference reads only phi(i, j, k, .)); prob_lo and dx are the
//    level's; time is 0 (no plugin of the reference reads it there);
//  * species arrays of the thermo functions are `const Real *`.
//
// Required members of Case:
//   static constexpr int PLUGIN_ID;            mf_plugin_params::kind (>= 100)
//   static constexpr int NUM_SPECIES, NUM_GAS_SPECIES;
//   static constexpr bool HAS_CHEMISTRY;       compute_spec_source is not zero
//   static constexpr bool HAS_FLUID_SOURCE;    compute_fluid_source is not zero
//   static constexpr bool GAMMA_IGNORES_PRESSURE;  gamma does not depend on p (required:
//                                              the WENO5 sweeps stage one gamma per cell)
//   static constexpr bool HAS_BG_GAS;          a background gas not tracked as a
//                                              species (mflo.using_bg_inertgas)
//   static void init_species(Params &);        species.cpp init(): nspec, ngas,
//                                              molwt, advect_flag (host)
//   MFP_HD static Real get_r_from_c(const Real *spec, Real bgasconc, const Params &);
//   MFP_HD static Real get_bgasconc_from_rc(Real rho, const Real *spec, const Params &);
//   MFP_HD static Real get_t_from_rpc(Real rho, Real pres, const Real *spec, const Params &);
//   MFP_HD static Real get_p_from_rec(Real rhoe, Real rho, const Real *spec, const Params &);
//   MFP_HD static Real get_gama_from_rpc(Real rho, Real pres, const Real *spec, const Params &);
//   MFP_HD static Real get_e_from_rpc(Real rho, Real pres, const Real *spec, const Params &);
//   MFP_HD static Real get_solid_t_from_rhoe(Real rhoe, const Params &);
//   MFP_HD static void compute_fluid_transport(int i, int j, int k, Array4<Real> const &phi,
//       Array4<Real> const &transpcoeffs, GpuArray<Real, 3> prob_lo, GpuArray<Real, 3> dx, Real time,
//       const Params &);
//   MFP_HD static void compute_fluid_source(...same..., Array4<Real> const &source, ...);
//   MFP_HD static void compute_spec_dcoeff(...same..., Array4<Real> const &dcoeff, ...);
//   MFP_HD static void compute_spec_source(...same..., Array4<Real> const &specsource, ...);
//       (specsource holds TOTAL_NVARS components, zeroed before the call;
//        the species sources go to FLO_NVARS + sp)
//   MFP_HD static void externalbc(const Real x[3], const Real s_int[], Real s_ext[], int idir, int sgn,
//       Real time, GeometryData const &geomdata, const Params &);
#pragma once
#include <stdint.h>

#include "mesoflow_amd.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MFP_HD __host__ __device__ __forceinline__
#else
#define MFP_HD inline
#endif

namespace mfp {

using Real = double;
using Params = mf_plugin_params;

// component indices of Source/globalDefines.H:19-32 under the reference's names
enum : int {
    RHO_INDX = 0, RHOU_INDX = 1, RHOV_INDX = 2, RHOW_INDX = 3, RHOE_INDX = 4, DENS_INDX = 5, VELX_INDX = 6,
    VELY_INDX = 7, VELZ_INDX = 8, PRES_INDX = 9, TEMP_INDX = 10, VFRAC_INDX = 11, FLO_NVARS = 12, NCVARS = 5
};
enum : int { VISC_INDX = 0, THCOND_INDX = 1 };
enum : int { XDIR = 0, YDIR = 1, ZDIR = 2 };

// Source/mflo_constants.H values under their names
constexpr Real zeroval = 0.0, one = 1.0, two = 2.0, three = 3.0, four = 4.0, five = 5.0, six = 6.0;
constexpr Real half = 0.5, tenth = 0.1, fourth = 0.25, eighth = 0.125;
constexpr Real RU = 8.31446261815324;    // J/K/mol
constexpr Real P_NTP = 101325.0;         // Pa

// amrex::Array4: (i, j, k, n) -> p[(i - begin.x) + (j - begin.y) jstride + (k - begin.z) kstride + n nstride]
template <class T>
struct Array4 {
    T *p;
    int64_t jstride, kstride, nstride;
    int bx, by, bz;
    MFP_HD T &operator()(int i, int j, int k, int n = 0) const {
        return p[(int64_t)(i - bx) + (int64_t)(j - by) * jstride + (int64_t)(k - bz) * kstride + (int64_t)n * nstride];
    }
};
// a view of one cell's components v[0..] at (i, j, k)
template <class T>
MFP_HD Array4<T> cell_view(T *v, int i, int j, int k) {
    Array4<T> a;
    a.p = v;
    a.jstride = 0;
    a.kstride = 0;
    a.nstride = 1;
    a.bx = i;
    a.by = j;
    a.bz = k;
    return a;
}

template <class T, int N>
struct GpuArray {
    T arr[N];
    MFP_HD const T &operator[](int q) const { return arr[q]; }
    MFP_HD T &operator[](int q) { return arr[q]; }
};

// amrex::GeometryData: problem bounds and cell sizes of the level
struct GeometryData {
    Real prob_lo[3], prob_hi[3], dx[3];
    MFP_HD const Real *ProbLo() const { return prob_lo; }
    MFP_HD const Real *ProbHi() const { return prob_hi; }
    MFP_HD const Real *CellSize() const { return dx; }
    MFP_HD Real ProbLo(int d) const { return prob_lo[d]; }
    MFP_HD Real ProbHi(int d) const { return prob_hi[d]; }
    MFP_HD Real CellSize(int d) const { return dx[d]; }
};

}  // namespace mfp
