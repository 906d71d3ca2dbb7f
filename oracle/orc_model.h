/*
 * TEST INFRASTRUCTURE ONLY — the oracle's array-level model description.
 *
 * Shared by the CPU restatement (cxg_oracle.cpp) and the reference shim
 * (ref_shim.cpp, which drives the reference's own calculixpp_core). Only tests,
 * the smoke check and the CPU-baseline leg of bench.py use anything under oracle/;
 * the product path (calculixpp_amd, libcxg.so) never links or calls it.
 *
 * Conventions follow the reference:
 *  - element type codes equal cxpp::ElementType's enumerator order
 *    (include/calculixpp/core/element.hpp:13-22): C3D4=0, C3D10=1, C3D8=2,
 *    C3D8R=3, C3D20=4, C3D20R=5, C3D6=6, C3D15=7;
 *  - connectivity holds node *indices* (0-based, insertion order), in
 *    CalculiX node order; the reference maps ids to indices through a hash
 *    (src/core/mesh.cpp:29-32), here the caller passes the indices directly;
 *  - global DOF g = 3*node_index + comp (comp 0..2), types.hpp:13;
 *  - Voigt order xx,yy,zz,xy,xz,yz with engineering shear.
 */
#ifndef CXG_ORC_MODEL_H
#define CXG_ORC_MODEL_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_C3D4 = 0, ORC_C3D10 = 1, ORC_C3D8 = 2, ORC_C3D8R = 3, ORC_C3D20 = 4,
       ORC_C3D20R = 5, ORC_C3D6 = 6, ORC_C3D15 = 7 };

/* Hardening kind, cxpp::Plastic::Hardening (core/material.hpp:27-28). */
enum { ORC_HARD_ISO = 0, ORC_HARD_KIN = 1, ORC_HARD_COMBINED = 2 };

#define ORC_MAX_TABLE 16

/* One J2 material (core/material.hpp:24-44 + ElasticIso). */
typedef struct {
  double E, nu;
  int32_t hardening;            /* ORC_HARD_* */
  int32_t n_table;              /* points in the *PLASTIC table (>=1) */
  double eqplastic[ORC_MAX_TABLE];
  double yield[ORC_MAX_TABLE];
  double kinematic_modulus;
} orc_j2;

/* Per Gauss point history, 26 doubles (fem/material_model.hpp:38-48):
 *   [0..5]   committed plastic strain (engineering shear)
 *   [6]      committed equivalent plastic strain
 *   [7..12]  committed back stress
 *   [13..18] trial plastic strain
 *   [19]     trial equivalent plastic strain
 *   [20..25] trial back stress                                                   */
#define ORC_STATE 26

typedef struct {
  int32_t n_nodes;
  const double* coords;   /* n_nodes*3 */
  int32_t n_elem;
  const int32_t* etype;   /* n_elem */
  const int64_t* eptr;    /* n_elem+1 offsets into conn */
  const int32_t* conn;    /* node indices */
  const double* E;        /* n_elem */
  const double* nu;       /* n_elem */
  const int32_t* j2_id;   /* n_elem, -1 = elastic; may be NULL (all elastic) */
  int32_t n_j2;
  const orc_j2* j2;       /* n_j2 materials */
  const uint8_t* active;  /* n_elem, 0 = removed (*MODEL CHANGE); may be NULL */
  int32_t n_spc;
  const int32_t* spc_node;   /* node index */
  const int32_t* spc_comp;   /* 1..3 */
  const double* spc_value;
  int32_t n_cload;
  const int32_t* cload_node; /* node index */
  const int32_t* cload_comp; /* 1..3 */
  const double* cload_value;
  const double* rho;         /* n_elem density (*DENSITY); may be NULL (no mass) */
} orc_model;

#ifdef __cplusplus
}
#endif
#endif
