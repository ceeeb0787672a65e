/*
 * C ABI of the MI355X FastNoise2 backend: the 45 entry points of FastNoise2's
 * include/FastNoise/FastNoise_C.h with identical names, signatures and return
 * conventions, so the existing C#/Rust/Java/Unreal bindings load this library
 * unchanged. Each declaration names the reference declaration it replaces
 * (FastNoise_C.h:line). Nodes are opaque handles; fnNew* return a handle
 * holding one reference, fnDeleteNodeRef drops it. Generation writes the
 * caller's host buffer; the work runs on the GPU.
 *
 * maxFeatureSet: 0 selects the default backend; any value at or above
 * FastSIMD::FeatureSet::GPU (0x1000) selects the GPU backend; values in
 * between request a CPU SIMD level, which this library does not provide, so
 * creation returns NULL.
 */
#ifndef FASTNOISE_C_H
#define FASTNOISE_C_H

#include "Utility/Export.h"

#ifdef __cplusplus
extern "C" {
#else
#include <stdbool.h>
#endif

/* FastNoise_C.h:27 */
FASTNOISE_API void* fnNewFromEncodedNodeTree( const char* encodedString, unsigned /*FastSIMD::FeatureSet*/ maxFeatureSet );
/* FastNoise_C.h:32 */
FASTNOISE_API void fnDeleteNodeRef( void* node );
/* FastNoise_C.h:38 */
FASTNOISE_API unsigned fnGetActiveFeatureSet( const void* node );
/* FastNoise_C.h:44 */
FASTNOISE_API int fnGetMetadataID( const void* node );

/* FastNoise_C.h:66 */
FASTNOISE_API void fnGenUniformGrid2D( const void* node, float* noiseOut,
    float xOffset, float yOffset,
    int xCount, int yCount,
    float xStepSize, float yStepSize,
    int seed, float* outputMinMax /*nullable*/ );

/* FastNoise_C.h:99 */
FASTNOISE_API void fnGenUniformGrid3D( const void* node, float* noiseOut,
    float xOffset, float yOffset, float zOffset,
    int xCount, int yCount, int zCount,
    float xStepSize, float yStepSize, float zStepSize,
    int seed, float* outputMinMax /*nullable*/ );

/* FastNoise_C.h:134 */
FASTNOISE_API void fnGenUniformGrid4D( const void* node, float* noiseOut,
    float xOffset, float yOffset, float zOffset, float wOffset,
    int xCount, int yCount, int zCount, int wCount,
    float xStepSize, float yStepSize, float zStepSize, float wStepSize,
    int seed, float* outputMinMax /*nullable*/ );

/* FastNoise_C.h:164 */
FASTNOISE_API void fnGenPositionArray2D( const void* node, float* noiseOut, int count,
    const float* xPosArray, const float* yPosArray,
    float xOffset, float yOffset,
    int seed, float* outputMinMax /*nullable*/ );

/* FastNoise_C.h:195 */
FASTNOISE_API void fnGenPositionArray3D( const void* node, float* noiseOut, int count,
    const float* xPosArray, const float* yPosArray, const float* zPosArray,
    float xOffset, float yOffset, float zOffset,
    int seed, float* outputMinMax /*nullable*/ );

/* FastNoise_C.h:228 */
FASTNOISE_API void fnGenPositionArray4D( const void* node, float* noiseOut, int count,
    const float* xPosArray, const float* yPosArray, const float* zPosArray, const float* wPosArray,
    float xOffset, float yOffset, float zOffset, float wOffset,
    int seed, float* outputMinMax /*nullable*/ );

/* FastNoise_C.h:250 */
FASTNOISE_API void fnGenTileable2D( const void* node, float* noiseOut,
    int xSize, int ySize,
    float xStepSize, float yStepSize,
    int seed, float* outputMinMax /*nullable*/ );

/* FastNoise_C.h:267 */
FASTNOISE_API float fnGenSingle2D( const void* node, float x, float y, int seed );
/* FastNoise_C.h:282 */
FASTNOISE_API float fnGenSingle3D( const void* node, float x, float y, float z, int seed );
/* FastNoise_C.h:298 */
FASTNOISE_API float fnGenSingle4D( const void* node, float x, float y, float z, float w, int seed );

/* FastNoise_C.h:307 */
FASTNOISE_API int fnGetMetadataCount();
/* FastNoise_C.h:318 */
FASTNOISE_API const char* fnGetMetadataName( int id );
/* FastNoise_C.h:332 */
FASTNOISE_API void* fnNewFromMetadata( int id, unsigned /*FastSIMD::FeatureSet*/ maxFeatureSet );

/* FastNoise_C.h:343 */
FASTNOISE_API int fnGetMetadataVariableCount( int id );
/* FastNoise_C.h:353 */
FASTNOISE_API const char* fnGetMetadataVariableName( int id, int variableIndex );
/* FastNoise_C.h:368 */
FASTNOISE_API int fnGetMetadataVariableType( int id, int variableIndex );
/* FastNoise_C.h:380 */
FASTNOISE_API int fnGetMetadataVariableDimensionIdx( int id, int variableIndex );
/* FastNoise_C.h:391 */
FASTNOISE_API int fnGetMetadataEnumCount( int id, int variableIndex );
/* FastNoise_C.h:403 */
FASTNOISE_API const char* fnGetMetadataEnumName( int id, int variableIndex, int enumIndex );
/* FastNoise_C.h:414 */
FASTNOISE_API bool fnSetVariableFloat( void* node, int variableIndex, float value );
/* FastNoise_C.h:427 */
FASTNOISE_API bool fnSetVariableIntEnum( void* node, int variableIndex, int value );

/* FastNoise_C.h:438 */
FASTNOISE_API int fnGetMetadataNodeLookupCount( int id );
/* FastNoise_C.h:448 */
FASTNOISE_API const char* fnGetMetadataNodeLookupName( int id, int nodeLookupIndex );
/* FastNoise_C.h:460 */
FASTNOISE_API int fnGetMetadataNodeLookupDimensionIdx( int id, int nodeLookupIndex );
/* FastNoise_C.h:472 */
FASTNOISE_API bool fnSetNodeLookup( void* node, int nodeLookupIndex, const void* nodeLookup );

/* FastNoise_C.h:486 */
FASTNOISE_API int fnGetMetadataHybridCount( int id );
/* FastNoise_C.h:496 */
FASTNOISE_API const char* fnGetMetadataHybridName( int id, int hybridIndex );
/* FastNoise_C.h:508 */
FASTNOISE_API int fnGetMetadataHybridDimensionIdx( int id, int hybridIndex );
/* FastNoise_C.h:521 */
FASTNOISE_API bool fnSetHybridNodeLookup( void* node, int hybridIndex, const void* nodeLookup );
/* FastNoise_C.h:534 */
FASTNOISE_API bool fnSetHybridFloat( void* node, int hybridIndex, float value );

/* FastNoise_C.h:545 */
FASTNOISE_API const char* fnGetMetadataDescription( int id );
/* FastNoise_C.h:555 */
FASTNOISE_API int         fnGetMetadataGroupCount( int id );
/* FastNoise_C.h:567 */
FASTNOISE_API const char* fnGetMetadataGroupName( int id, int groupIndex );
/* FastNoise_C.h:578 */
FASTNOISE_API const char* fnGetMetadataVariableDescription( int id, int variableIndex );
/* FastNoise_C.h:589 */
FASTNOISE_API float       fnGetMetadataVariableDefaultFloat( int id, int variableIndex );
/* FastNoise_C.h:600 */
FASTNOISE_API int         fnGetMetadataVariableDefaultIntEnum( int id, int variableIndex );
/* FastNoise_C.h:611 */
FASTNOISE_API float       fnGetMetadataVariableMinFloat( int id, int variableIndex );
/* FastNoise_C.h:622 */
FASTNOISE_API float       fnGetMetadataVariableMaxFloat( int id, int variableIndex );
/* FastNoise_C.h:633 */
FASTNOISE_API const char* fnGetMetadataNodeLookupDescription( int id, int nodeLookupIndex );
/* FastNoise_C.h:644 */
FASTNOISE_API const char* fnGetMetadataHybridDescription( int id, int hybridIndex );
/* FastNoise_C.h:651 */
FASTNOISE_API float       fnGetMetadataHybridDefault( int id, int hybridIndex );

#ifdef __cplusplus
}
#endif

#endif
