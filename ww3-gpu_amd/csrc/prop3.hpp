// ww3-gpu_amd/csrc/prop3.hpp
// The PR3/UQ propagation on gfx950: W3XYP3 + W3QCK3 over every spectral component of a step,
// and W3KTP3 + W3QCK1/W3QCK2 over the sea points of one intra-spectral part.
//
// Reference: kokkos/src/ww_kokkos/prop_xyp3.hpp:156-584, prop_qck3.hpp:58-214,
// prop_ktp3.hpp:143-353, prop_qck1.hpp:54-177, prop_qck2.hpp:60-193; the batch entries
// kokkos/src/fortran_iface/prop3_shim.cpp:470-750.
//
// W3XYP3 work mapping.  Every component is one plane FIELD(1-NY : NY*(NX+2)) in HBM, the grid
// row index fastest, so a sweep over the cells of one component is a coalesced stream; a
// launch covers all components of the call at once (blockIdx.y = component).  A sub-step is
// five passes over the planes: the garden-sprinkler averaging (restated as an ordered gather:
// each cell adds the contributions of its nine neighbours in the order the reference's serial
// scatter stores them, so every sum keeps its bits), then per sweep the QCK3 fluxes and the
// update, then the back-transformation and the boundary input.  Components with fewer CFL
// sub-steps drop out of the later launches.
//
// W3KTP3 work mapping.  One 64-lane workgroup per sea point; the point's rotated spectrum
// VQ(-NK-1 : NK2*(NTH+2)) and the velocity and flux arrays live in LDS; every W3QCK1/W3QCK2
// loop is split over the lanes (its map entries are distinct cells).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ww3gpu {
namespace prop {

/// One component of a W3XYP3 call: W3XYP3's per-component scalars (prop_xyp3.hpp:200-275),
/// computed on the host in the reference's operation order.
struct Comp {
  int isp, ik, ith, ntloc;
  float rntloc, dtrad, fthx, fthy, fcgx, fcgy, ecosi, esini;
};

/// A sea point's averaging stencil for one component (prop_xyp3.hpp:983-1059 and the RD
/// factors of :1160-1163): RD1, MIN(RDI2*CG, 1), RD3, MIN(CG*RDI4, 1).
struct Rec {
  float rd1, rd2, rd3, rd4;
};

/// Where a sea point's nine averaging stores go, grouped by target direction (the index of the
/// target's offset in {0, 1, -1, NY, -NY, NY+1, NY-1, -NY+1, -NY-1}), each group in the order the
/// point's loop makes its stores (centre first; the central loop's +1..+4, -1..-4 or the
/// near-coast loop's +1, -1, .. -4, each already sent back to the centre where the reference
/// redirects it).  Direction d = 1..8 uses bits 6(d-1) .. 6(d-1)+5: the number of stores there
/// (0, 1 or 2: the eight stores hold four offsets and their negatives, so a direction gets at
/// most two) and the factor set - 1 of each (2 bits).  Direction 0 receives the centre store
/// (factor set 0) and then the near-coast stores j = 1..8 sent back, a mask in bits 48..55
/// (store j at bit 47 + j).  Bit 62 marks a near-coast point.
constexpr int kCentreShift = 48;
constexpr int kNearCoastBit = 62;

/// The grid of a W3XYP3 call, device pointers.  Cell IXY = IY + (IX-1)*NY, 1-based.
struct XyGrid {
  int nx, ny, nsea, nk, nth, nspec, nxy, nq, nf, cgld;
  int global, flcur, ncent, flbpi;
  float dtg, dsec_bpi0, dsec_bpin;
  const int* mapsf;    ///< MAPSF(NSEA,3)
  const int* mapsta;   ///< MAPSTA(NY*NX)
  const int* maptrn;   ///< MAPTRN(NY*NX), 0/1
  const int* cellsea;  ///< [NXY+1]: ISEA of cell IXY, 0 for none
  const int* pos;      ///< [NSEA+1]: position of ISEA in MAPCXY (1..NSEA)
  /// [NQ]: for target cell t = k+1-NY, the neighbour directions (0..8, 4 bits each, low first,
  /// 0xF ends) whose source cell scatters in the averaging, in the order its stores happen.
  const uint64_t* order;
  const unsigned char* clsx;  ///< [NXY+1]: 1 central, 2 above, 3 below (MAPX2 sections), 0
  const unsigned char* clsy;  ///< [NXY+1]: the same for MAPY2
  const unsigned char* act;   ///< [NXY+1]: listed in MAPAXY
  const int* bnd;             ///< [NXY+1]: the last IBI whose point is cell IXY, 0 for none
  const float *clats, *dpdx, *dpdy, *dqdx, *dqdy, *gsqrt, *atrnx, *atrny, *cg, *cx, *cy;
  const int* isbpi;
  const float *bbpi0, *bbpin;
};

/// The device arrays of one W3XYP3 call.
struct XyState {
  int nlist;
  const Comp* comps;  ///< [nlist]
  Rec* rec;           ///< [nlist][NXY+1], by the source's cell
  uint64_t* tw;       ///< [nlist][NXY+1]: the store targets (kNearCoastBit)
  float* q[2];        ///< [nlist][NQ] each: FIELD, ping-pong across sub-steps
  float* fla;         ///< [nlist][NF]: FLA(1-NY : NY*NX)
  /// [nlist][NF] each: VLCFLX(1 : NY*(NX+1)) with the extension column as W3QCK3's closure
  /// leaves it (column 1 on a closed grid, else 0), and VLCFLY (extension column 0)
  float* vlcx;
  float* vlcy;
};

hipError_t xyp3_gather(const XyGrid& g, const XyState& s, const float* va, hipStream_t st);
hipError_t xyp3_records(const XyGrid& g, const XyState& s, hipStream_t st);
/// One sub-step (averaging, the two sweeps in the order of this sub-step, back-transform and
/// boundary input) for every component with NTLOC >= itloc.
hipError_t xyp3_substep(const XyGrid& g, const XyState& s, int itloc, bool yfirst, bool flcx,
                        bool flcy, hipStream_t st);
/// Section 4 and W3SCAT: the components' results back into VA where MAPSTA /= 0.
hipError_t xyp3_store(const XyGrid& g, const XyState& s, float* va, hipStream_t st);

/// W3KTP3's grid (prop_ktp3.hpp:35-66), device pointers.
struct KtGrid {
  int nk, nth, nspec, nk2, flcth, flck, flcur;
  float dmin, ctmax, fachfa;
  const float *sig, *dsip, *ecos, *esin, *ec2, *esc, *es2;
  const int *mapwn, *mapth2, *mapwn2;
};

/// NTIMES W3KTP3 calls on each of npts points; VA is VA(NSPEC,0:NSEALM).
hipError_t ktp3_batch(const KtGrid& g, int itstep, int ntimes, float facth, float fack,
                      const float* cg, const float* wn, int npts, const int* jsea,
                      const int* isea, const float* pscal, float* va, float* cflthmax,
                      float* cflkmax, hipStream_t st);
/// LDS bytes of one W3KTP3 workgroup.
size_t ktp3_lds_bytes(int nk, int nth);

}  // namespace prop
}  // namespace ww3gpu
