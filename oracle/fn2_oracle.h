/*
 * Test infrastructure only: C interface of the CPU oracle for FastNoise2's
 * generator path.
 *
 * The oracle is a CPU restatement of the reference algorithm (node math in
 * the include/FastNoise/Generators .inl files, drivers in Generator.inl, wire format
 * in src/FastNoise/Metadata.cpp and Base64.h of FastNoise2 v1.1.1). It is the
 * checker for the parity tests and the host baseline of bench.py; the product
 * library never links it.
 *
 * Modes (bit flags):
 *   FN2O_UNFUSED  every FS::FMulAdd/FNMulAdd/FMulSub site evaluated as a
 *                 separately rounded multiply and add instead of one fma
 *   FN2O_AVX2     use the AVX2 branch of the helpers whose AVX2 branch rounds
 *                 differently from the generic one (GetGradientDotPerlin 2D,
 *                 Utils.inl:174-180); otherwise the generic (Scalar) branch
 */
#ifndef FN2_ORACLE_H
#define FN2_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define FN2O_UNFUSED 1
#define FN2O_AVX2 2

/* Decode an encoded node tree (Metadata.cpp:499-506). Returns NULL on failure. */
void* fn2o_new_from_encoded( const char* encoded );
void fn2o_delete( void* tree );
/* Node id of the root (registration order, FastSIMD_Build.inl:86-143). */
int fn2o_root_id( const void* tree );

/* Drivers (Generator.inl:79-433). lanes: 1 or 16; threads >= 1.
 * minmax may be NULL. Return 0 on success, -1 on bad arguments. */
int fn2o_gen_grid2d( const void* tree, float* out, float xo, float yo, int nx, int ny, float sx, float sy, int seed, int mode, int lanes, int threads, float* minmax );
int fn2o_gen_grid3d( const void* tree, float* out, float xo, float yo, float zo, int nx, int ny, int nz, float sx, float sy, float sz, int seed, int mode, int lanes, int threads, float* minmax );
int fn2o_gen_grid4d( const void* tree, float* out, float xo, float yo, float zo, float wo, int nx, int ny, int nz, int nw, float sx, float sy, float sz, float sw, int seed, int mode, int lanes, int threads, float* minmax );
int fn2o_gen_pos2d( const void* tree, float* out, int count, const float* xs, const float* ys, float xo, float yo, int seed, int mode, int lanes, int threads, float* minmax );
int fn2o_gen_pos3d( const void* tree, float* out, int count, const float* xs, const float* ys, const float* zs, float xo, float yo, float zo, int seed, int mode, int lanes, int threads, float* minmax );
int fn2o_gen_pos4d( const void* tree, float* out, int count, const float* xs, const float* ys, const float* zs, const float* ws, float xo, float yo, float zo, float wo, int seed, int mode, int lanes, int threads, float* minmax );
int fn2o_gen_tileable2d( const void* tree, float* out, int nx, int ny, float sx, float sy, int seed, int mode, int lanes, int threads, float* minmax );
float fn2o_gen_single2d( const void* tree, float x, float y, int seed, int mode );
float fn2o_gen_single3d( const void* tree, float x, float y, float z, int seed, int mode );
float fn2o_gen_single4d( const void* tree, float x, float y, float z, float w, int seed, int mode );

/* Shared-math probes for the closed-form tests (Utils.inl). */
int fn2o_hash_primes( int seed, int n, const int* primed );      /* HashPrimes */
int fn2o_hash_primes_hb( int seed, int n, const int* primed );   /* HashPrimesHB */
float fn2o_value_coord( int seed, int n, const int* primed );    /* GetValueCoord */
float fn2o_grad_dot( int dims, int hash, const float* f, int mode ); /* GetGradientDotPerlin 2D/4D, Common 3D */
float fn2o_interp( int kind, float t, int mode );                /* 0 Hermite, 1 Quintic */
/* One vertex's vector contribution of the simplex-type warps
 * (DomainWarpSimplex.inl:6-724): scheme 0 OrthogonalGradientMatrix,
 * 1 GradientOuterProduct; v[0..dims) accumulates mult * (vector for hash
 * applied to offset f). */
void fn2o_warp_vec( int dims, int scheme, int hash, const float* f, float mult, float* v, int mode );

/* Base64 with the '@' run-length extension (Base64.h). Returns decoded length or -1. */
int fn2o_base64_decode( const char* in, unsigned char* out, int cap );

#ifdef __cplusplus
}
#endif
#endif
