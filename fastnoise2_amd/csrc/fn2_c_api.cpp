// FastNoise2's C ABI (include/FastNoise/FastNoise_C.h of the reference) over
// the MI355X backend. Return conventions follow src/FastNoise/FastNoise_C.cpp
// of the reference: "INVALID ..." strings, -1 counts and 0 defaults for bad
// ids or indices, false from setters with an out-of-range index. Node ids are
// narrowed to Metadata::node_id (uint8_t) before lookup, as there.
#include <FastNoise/FastNoise_C.h>
#include <FastNoise/FastNoise.h>
#include <FastNoise/Metadata.h>
#include "fn2_backend.h"

using FastNoise::Generator;
using FastNoise::Metadata;

namespace
{
    const Generator* G( const void* p ) { return static_cast<const Generator*>( p ); }
    Generator* G( void* p ) { return static_cast<Generator*>( p ); }

    const Metadata* Meta( int id ) { return Metadata::GetFromId( (Metadata::node_id)id ); }

    void Store( float* mm, const FastNoise::OutputMinMax& r )
    {
        if( mm )
        {
            mm[0] = r.min;
            mm[1] = r.max;
        }
    }

    bool NodeOk( const void* node )
    {
        if( node ) return true;
        fn2::SetError( "null node handle" );
        return false;
    }

    template<typename V>
    const char* NameOr( const Metadata* m, const V& members, int index, const char* bad )
    {
        if( !m ) return "INVALID NODE ID";
        if( (size_t)index < members( m ).size() ) return members( m )[index].name;
        return bad;
    }
}

extern "C" {

void* fnNewFromEncodedNodeTree( const char* encodedString, unsigned maxFeatureSet )
{
    if( FastNoise::SmartNode<> node = FastNoise::NewFromEncodedNodeTree( encodedString, (FastSIMD::FeatureSet)maxFeatureSet ) )
    {
        FastNoise::Internal::BumpNodeRefences( node.get(), true );
        return node.get();
    }
    return nullptr;
}

void fnDeleteNodeRef( void* node )
{
    if( node ) FastNoise::Internal::BumpNodeRefences( G( node ), false );
}

unsigned fnGetActiveFeatureSet( const void* node )
{
    return node ? (unsigned)G( node )->GetActiveFeatureSet() : 0u;
}

int fnGetMetadataID( const void* node )
{
    return node ? G( node )->GetMetadata().id : -1;
}

void fnGenUniformGrid2D( const void* node, float* noiseOut, float xOffset, float yOffset, int xCount, int yCount, float xStepSize, float yStepSize, int seed, float* outputMinMax )
{
    if( !NodeOk( node ) ) return;
    Store( outputMinMax, G( node )->GenUniformGrid2D( noiseOut, xOffset, yOffset, xCount, yCount, xStepSize, yStepSize, seed ) );
}

void fnGenUniformGrid3D( const void* node, float* noiseOut, float xOffset, float yOffset, float zOffset, int xCount, int yCount, int zCount, float xStepSize, float yStepSize, float zStepSize, int seed, float* outputMinMax )
{
    if( !NodeOk( node ) ) return;
    Store( outputMinMax, G( node )->GenUniformGrid3D( noiseOut, xOffset, yOffset, zOffset, xCount, yCount, zCount, xStepSize, yStepSize, zStepSize, seed ) );
}

void fnGenUniformGrid4D( const void* node, float* noiseOut, float xOffset, float yOffset, float zOffset, float wOffset, int xCount, int yCount, int zCount, int wCount, float xStepSize, float yStepSize, float zStepSize, float wStepSize, int seed, float* outputMinMax )
{
    if( !NodeOk( node ) ) return;
    Store( outputMinMax, G( node )->GenUniformGrid4D( noiseOut, xOffset, yOffset, zOffset, wOffset, xCount, yCount, zCount, wCount, xStepSize, yStepSize, zStepSize, wStepSize, seed ) );
}

void fnGenPositionArray2D( const void* node, float* noiseOut, int count, const float* xPosArray, const float* yPosArray, float xOffset, float yOffset, int seed, float* outputMinMax )
{
    if( !NodeOk( node ) ) return;
    Store( outputMinMax, G( node )->GenPositionArray2D( noiseOut, count, xPosArray, yPosArray, xOffset, yOffset, seed ) );
}

void fnGenPositionArray3D( const void* node, float* noiseOut, int count, const float* xPosArray, const float* yPosArray, const float* zPosArray, float xOffset, float yOffset, float zOffset, int seed, float* outputMinMax )
{
    if( !NodeOk( node ) ) return;
    Store( outputMinMax, G( node )->GenPositionArray3D( noiseOut, count, xPosArray, yPosArray, zPosArray, xOffset, yOffset, zOffset, seed ) );
}

void fnGenPositionArray4D( const void* node, float* noiseOut, int count, const float* xPosArray, const float* yPosArray, const float* zPosArray, const float* wPosArray, float xOffset, float yOffset, float zOffset, float wOffset, int seed, float* outputMinMax )
{
    if( !NodeOk( node ) ) return;
    Store( outputMinMax, G( node )->GenPositionArray4D( noiseOut, count, xPosArray, yPosArray, zPosArray, wPosArray, xOffset, yOffset, zOffset, wOffset, seed ) );
}

void fnGenTileable2D( const void* node, float* noiseOut, int xSize, int ySize, float xStepSize, float yStepSize, int seed, float* outputMinMax )
{
    if( !NodeOk( node ) ) return;
    Store( outputMinMax, G( node )->GenTileable2D( noiseOut, xSize, ySize, xStepSize, yStepSize, seed ) );
}

float fnGenSingle2D( const void* node, float x, float y, int seed )
{
    if( !NodeOk( node ) ) return __builtin_nanf( "" );
    return G( node )->GenSingle2D( x, y, seed );
}

float fnGenSingle3D( const void* node, float x, float y, float z, int seed )
{
    if( !NodeOk( node ) ) return __builtin_nanf( "" );
    return G( node )->GenSingle3D( x, y, z, seed );
}

float fnGenSingle4D( const void* node, float x, float y, float z, float w, int seed )
{
    if( !NodeOk( node ) ) return __builtin_nanf( "" );
    return G( node )->GenSingle4D( x, y, z, w, seed );
}

int fnGetMetadataCount()
{
    return (int)Metadata::GetAll().size();
}

const char* fnGetMetadataName( int id )
{
    const Metadata* m = Meta( id );
    return m ? m->name : "INVALID NODE ID";
}

void* fnNewFromMetadata( int id, unsigned maxFeatureSet )
{
    if( const Metadata* m = Meta( id ) )
    {
        FastNoise::SmartNode<> node = m->CreateNode( (FastSIMD::FeatureSet)maxFeatureSet );
        if( !node ) return nullptr;
        FastNoise::Internal::BumpNodeRefences( node.get(), true );
        return node.get();
    }
    return nullptr;
}

int fnGetMetadataVariableCount( int id )
{
    const Metadata* m = Meta( id );
    return m ? (int)m->memberVariables.size() : -1;
}

const char* fnGetMetadataVariableName( int id, int variableIndex )
{
    return NameOr( Meta( id ), []( const Metadata* m ) -> const auto& { return m->memberVariables; }, variableIndex, "INVALID VARIABLE INDEX" );
}

int fnGetMetadataVariableType( int id, int variableIndex )
{
    const Metadata* m = Meta( id );
    if( m && (size_t)variableIndex < m->memberVariables.size() ) return (int)m->memberVariables[variableIndex].type;
    return -1;
}

int fnGetMetadataVariableDimensionIdx( int id, int variableIndex )
{
    const Metadata* m = Meta( id );
    if( m && (size_t)variableIndex < m->memberVariables.size() ) return m->memberVariables[variableIndex].dimensionIdx;
    return -1;
}

int fnGetMetadataEnumCount( int id, int variableIndex )
{
    const Metadata* m = Meta( id );
    if( m && (size_t)variableIndex < m->memberVariables.size() ) return (int)m->memberVariables[variableIndex].enumNames.size();
    return -1;
}

const char* fnGetMetadataEnumName( int id, int variableIndex, int enumIndex )
{
    const Metadata* m = Meta( id );
    if( !m ) return "INVALID NODE ID";
    if( (size_t)variableIndex >= m->memberVariables.size() ) return "INVALID VARIABLE INDEX";
    const auto& names = m->memberVariables[variableIndex].enumNames;
    if( (size_t)enumIndex >= names.size() ) return "INVALID ENUM INDEX";
    return names[enumIndex];
}

bool fnSetVariableFloat( void* node, int variableIndex, float value )
{
    if( !NodeOk( node ) ) return false;
    const Metadata& m = G( node )->GetMetadata();
    if( (size_t)variableIndex < m.memberVariables.size() ) return m.memberVariables[variableIndex].setFunc( G( node ), value );
    return false;
}

bool fnSetVariableIntEnum( void* node, int variableIndex, int value )
{
    if( !NodeOk( node ) ) return false;
    const Metadata& m = G( node )->GetMetadata();
    if( (size_t)variableIndex < m.memberVariables.size() ) return m.memberVariables[variableIndex].setFunc( G( node ), value );
    return false;
}

int fnGetMetadataNodeLookupCount( int id )
{
    const Metadata* m = Meta( id );
    return m ? (int)m->memberNodeLookups.size() : -1;
}

const char* fnGetMetadataNodeLookupName( int id, int nodeLookupIndex )
{
    return NameOr( Meta( id ), []( const Metadata* m ) -> const auto& { return m->memberNodeLookups; }, nodeLookupIndex, "INVALID NODE LOOKUP INDEX" );
}

int fnGetMetadataNodeLookupDimensionIdx( int id, int nodeLookupIndex )
{
    const Metadata* m = Meta( id );
    if( m && (size_t)nodeLookupIndex < m->memberNodeLookups.size() ) return m->memberNodeLookups[nodeLookupIndex].dimensionIdx;
    return -1;
}

bool fnSetNodeLookup( void* node, int nodeLookupIndex, const void* nodeLookup )
{
    if( !NodeOk( node ) ) return false;
    const Metadata& m = G( node )->GetMetadata();
    if( (size_t)nodeLookupIndex < m.memberNodeLookups.size() )
    {
        FastNoise::SmartNode<> lookup = FastNoise::Internal::ToSmartNode( nodeLookup );
        return m.memberNodeLookups[nodeLookupIndex].setFunc( G( node ), lookup );
    }
    return false;
}

int fnGetMetadataHybridCount( int id )
{
    const Metadata* m = Meta( id );
    return m ? (int)m->memberHybrids.size() : -1;
}

const char* fnGetMetadataHybridName( int id, int hybridIndex )
{
    return NameOr( Meta( id ), []( const Metadata* m ) -> const auto& { return m->memberHybrids; }, hybridIndex, "INVALID HYBRID INDEX" );
}

int fnGetMetadataHybridDimensionIdx( int id, int hybridIndex )
{
    const Metadata* m = Meta( id );
    if( m && (size_t)hybridIndex < m->memberHybrids.size() ) return m->memberHybrids[hybridIndex].dimensionIdx;
    return -1;
}

bool fnSetHybridNodeLookup( void* node, int hybridIndex, const void* nodeLookup )
{
    if( !NodeOk( node ) ) return false;
    const Metadata& m = G( node )->GetMetadata();
    if( (size_t)hybridIndex < m.memberHybrids.size() )
    {
        FastNoise::SmartNode<> lookup = FastNoise::Internal::ToSmartNode( nodeLookup );
        return m.memberHybrids[hybridIndex].setNodeFunc( G( node ), lookup );
    }
    return false;
}

bool fnSetHybridFloat( void* node, int hybridIndex, float value )
{
    if( !NodeOk( node ) ) return false;
    const Metadata& m = G( node )->GetMetadata();
    if( (size_t)hybridIndex < m.memberHybrids.size() ) return m.memberHybrids[hybridIndex].setValueFunc( G( node ), value );
    return false;
}

const char* fnGetMetadataDescription( int id )
{
    const Metadata* m = Meta( id );
    return m ? m->description : "INVALID NODE ID";
}

int fnGetMetadataGroupCount( int id )
{
    const Metadata* m = Meta( id );
    return m ? (int)m->groups.size() : -1;
}

const char* fnGetMetadataGroupName( int id, int groupIndex )
{
    const Metadata* m = Meta( id );
    if( !m ) return "INVALID NODE ID";
    if( (size_t)groupIndex < m->groups.size() ) return m->groups[groupIndex];
    return "INVALID GROUP INDEX";
}

const char* fnGetMetadataVariableDescription( int id, int variableIndex )
{
    const Metadata* m = Meta( id );
    if( !m ) return "INVALID NODE ID";
    if( (size_t)variableIndex < m->memberVariables.size() ) return m->memberVariables[variableIndex].description;
    return "INVALID VARIABLE INDEX";
}

float fnGetMetadataVariableDefaultFloat( int id, int variableIndex )
{
    const Metadata* m = Meta( id );
    if( m && (size_t)variableIndex < m->memberVariables.size() ) return m->memberVariables[variableIndex].valueDefault.f;
    return 0;
}

int fnGetMetadataVariableDefaultIntEnum( int id, int variableIndex )
{
    const Metadata* m = Meta( id );
    if( m && (size_t)variableIndex < m->memberVariables.size() ) return m->memberVariables[variableIndex].valueDefault.i;
    return 0;
}

float fnGetMetadataVariableMinFloat( int id, int variableIndex )
{
    const Metadata* m = Meta( id );
    if( m && (size_t)variableIndex < m->memberVariables.size() ) return m->memberVariables[variableIndex].valueMin.f;
    return 0;
}

float fnGetMetadataVariableMaxFloat( int id, int variableIndex )
{
    const Metadata* m = Meta( id );
    if( m && (size_t)variableIndex < m->memberVariables.size() ) return m->memberVariables[variableIndex].valueMax.f;
    return 0;
}

const char* fnGetMetadataNodeLookupDescription( int id, int nodeLookupIndex )
{
    const Metadata* m = Meta( id );
    if( !m ) return "INVALID NODE ID";
    if( (size_t)nodeLookupIndex < m->memberNodeLookups.size() ) return m->memberNodeLookups[nodeLookupIndex].description;
    return "INVALID NODE LOOKUP INDEX";
}

const char* fnGetMetadataHybridDescription( int id, int hybridIndex )
{
    const Metadata* m = Meta( id );
    if( !m ) return "INVALID NODE ID";
    if( (size_t)hybridIndex < m->memberHybrids.size() ) return m->memberHybrids[hybridIndex].description;
    return "INVALID HYBRID INDEX";
}

float fnGetMetadataHybridDefault( int id, int hybridIndex )
{
    const Metadata* m = Meta( id );
    if( m && (size_t)hybridIndex < m->memberHybrids.size() ) return m->memberHybrids[hybridIndex].valueDefault;
    return 0;
}

}
