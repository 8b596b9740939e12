/*
 * kqed_assembly.h -- sparse (CSR) assembly maps of the QED kernel built from Dirac traces.
 */
#ifndef KQED_ASSEMBLY_H
#define KQED_ASSEMBLY_H

#ifdef __cplusplus
extern "C" {
#endif

#define KQED_MAP_MAXNNZ 8192

typedef struct {
  int rowptr[385];
  short col[KQED_MAP_MAXNNZ];
  double val[KQED_MAP_MAXNNZ];
  int nnz, kmax;
} kqed_map;

/* G^I, G^II, G^III as dense [3][384][64] (row = 64 rs + 16 mu + 4 nu + lambda,
 * col = 16 alpha + 4 beta + delta). */
void kqed_dense_maps(double *Gfull);
/* standard map (200 inputs), Tabd map (192 inputs) and the x=y map (Tabd inputs, mu<->lambda, sign) */
int kqed_build_maps(kqed_map *std_map, kqed_map *eq0_map, kqed_map *eq0_mulam_map);
/* the reference's 65536-entry eight-gamma trace table */
void kqed_fill_g8(int *G8);
/* 1/4 Tr of n <= 16 Euclidean gammas with the given indices (0 beyond 16) */
int kqed_quarter_trace(int n, const int *idx);

#ifdef __cplusplus
}
#endif
#endif
