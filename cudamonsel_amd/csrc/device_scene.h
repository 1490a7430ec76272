/*
 * device_scene.h -- device-resident scene layout of the transport engine.
 *
 * The host compiles a monsel_scene_desc (SI, FP64, the reference's object
 * graph) into these flat FP32 records in nm / eV units.  Every record is
 * plain data so that one hipMemcpy uploads the scene.
 */
#ifndef MONSEL_DEVICE_SCENE_H
#define MONSEL_DEVICE_SCENE_H

#include <stdint.h>

/* Scene records are read-only for the lifetime of a kernel, so device code
 * reads them through the constant address space: a load whose index is the
 * same on every lane of a wave then becomes one scalar load into SGPRs
 * instead of a vector load per lane.  (The host pass sees plain pointers of
 * the same size.) */
#if defined(__HIP_DEVICE_COMPILE__)
#define MONSEL_CS __attribute__((address_space(4)))
#else
#define MONSEL_CS
#endif

namespace monsel_dev {

constexpr int kMaxPrims = 8;      /* convex primitives per region shape */
constexpr int kMaxPlanes = 16;    /* planes per multi-plane primitive */
constexpr int kMaxElem = 8;       /* elements per material */
constexpr int kMaxMech = 6;       /* mechanisms per material scatter model */
constexpr int kMaxCore = 8;       /* core levels per material */
constexpr int kMottNE = 61;
constexpr int kMottNQ = 201;

enum PrimKind { PRIM_PLANES = 0, PRIM_CYLINDER = 1, PRIM_SPHERE = 2 };

/* One convex primitive.  Planes: inside is n.p - b <= 0. */
struct DPrim {
  int kind;
  int first_plane;
  int n_planes;
  float sdscale; /* planes: min over the planes of 1/|n| (distance bounds); 1 otherwise */
  float e0[3];   /* cylinder end0 / sphere centre (nm) */
  float ax[3];   /* cylinder unit axis */
  float len;     /* cylinder length (nm) */
  float r2;      /* radius^2 (nm^2) */
  float axv[3];  /* cylinder axis vector end1-end0 (nm), for point containment */
  float len2;    /* |axv|^2 */
  float4 pl[4];  /* the first planes of a multi-plane primitive again, inline (one
                  * fewer dependent load for the common primitives) */
};
constexpr int kInlinePlanes = 4;

/* Shapes of at most this many primitives (whose truth table is one 32-bit
 * word) get the sharper safety bound of shape_safety. */
constexpr int kMaxCstPrims = 5;

/* A region shape: a boolean function of up to kMaxPrims primitives, stored
 * as a 256-entry truth table indexed by the primitives' inside bits. */
struct DShape {
  int first_prim;
  int n_prims;
  int is_sphere; /* the chamber sphere: a single primitive */
  int pad;
  uint32_t truth[8];
};

struct DRegion {
  int parent;
  int shape;
  int msm;
  int first_child;
  int n_children;
  int depth;
  int pad[2];
  DShape sh;    /* its shape's record again, inline (one fewer dependent load) */
};

/* JoyLuoNieminenCSD (eV, nm). */
struct DCSD {
  int active;
  int nce;
  float breakE;  /* eV */
  float gamma;   /* 1/(nm eV^1.5) */
  float minE;    /* eV (barrier height) */
  float pad;
  float recipJ[kMaxElem]; /* 1/eV */
  float beta[kMaxElem];
  float coef[kMaxElem];   /* eV^2/nm */
};

/* One element of an elastic (SelectableElasticSM) mechanism. */
struct DElem {
  int kind;     /* MONSEL_MECH_*_ELASTIC */
  int mott;     /* table index or -1 */
  int method;
  int Z;
  float natoms; /* scalefactor * density * N_A, atoms / nm^3 */
  float extrap; /* eV, NISTMottRS extrapolateBelowEnergy */
  float mottAtMin; /* nm^2 */
  float sfBrowning;
  float zp17, zp2, zp3, z;
  float rz13;   /* Z^(1/3) */
  float rz23;   /* Z^(2/3) */
  float pad[2];
};

struct DMech {
  int kind;
  int elem0;
  int nelem;
  int method;
  int table[4];
  float p[24];
  DCSD csd; /* FittedInelSM's slowing-down algorithm */
  double q[24]; /* tabulated mechanism: the TP_* slots in SI units (J) for the FP64 scatter */
};

/* Tabulated-mechanism parameter slots in DMech::p (eV) and DMech::q (J). */
enum {
  TP_OFFSET = 0, TP_MINEGEN = 1, TP_RATEMULT = 2, TP_E0DISP = 3, TP_CB = 4, TP_BANDGAP = 5,
  TP_GAP = 6, TP_OFFSETFERMI = 7, TP_BEREF = 8, TP_EIDOM0 = 9, TP_EIDOM1 = 10, TP_IIMFPDOM0 = 11,
  TP_IIMFPDOM1 = 12, TP_SE0RANGE0 = 13, TP_SE0RANGE1 = 14, TP_NCORE = 15, TP_CORE0 = 16
};

struct DMSM {
  int is_null;
  int n_mech;
  int mech0;
  int barrier_kind;
  int elastic_k;   /* first elastic mechanism (index within this model), -1 if none */
  int pad[3];
  float minE;      /* eV, may be -inf */
  float cb;        /* eV, material conduction-band bottom */
  float lambdaFactor; /* ExpQMBarrierSM lambda factor, 1/sqrt(eV) */
  float lambda;    /* nm */
  DCSD csd;
};

/* Non-uniform table (NUTableInterpolation) in the file's own units. */
struct DNUTable {
  int dim;
  int n[4];
  int xoff[4];  /* offsets of the abscissae in nu_data */
  int voff;     /* offset of the values */
  int pad[2];
};

struct DevScene {
  int n_regions, n_shapes, n_prims, n_planes, n_msm, n_mech, n_elem, n_mott, n_nu;
  int has_tabulated;      /* some mechanism is TabulatedInelasticSM (selects the transport kernel) */
  uint32_t compat;
  float chamber_r;        /* nm */
  float chamber_c[3];     /* nm */
  const MONSEL_CS DRegion* regions;
  const MONSEL_CS int* children;
  const MONSEL_CS DShape* shapes;
  const MONSEL_CS DPrim* prims;
  const MONSEL_CS float4* planes;   /* (nx, ny, nz, b) */
  const MONSEL_CS DMSM* msms;
  const MONSEL_CS DMech* mechs;
  const MONSEL_CS DElem* elems;
  const MONSEL_CS float* mott_spwem; /* [n_mott][61] */
  const MONSEL_CS float* mott_x1;    /* [n_mott][61*201] */
  const MONSEL_CS DNUTable* nu;
  const MONSEL_CS float* nu_data;
  const MONSEL_CS double* nu_data64; /* the same values and abscissae in FP64 */
};

}  // namespace monsel_dev

#endif
