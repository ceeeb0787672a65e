// Public entry point of the MI355X FastNoise2 backend, API-compatible with
// FastNoise2's include/FastNoise/FastNoise.h: node classes, New<T>() and
// NewFromEncodedNodeTree().
#pragma once
#include <string>
#include "Utility/Config.h"

#include "Generators/BasicGenerators.h"
#include "Generators/Value.h"
#include "Generators/Perlin.h"
#include "Generators/Simplex.h"
#include "Generators/Cellular.h"
#include "Generators/Fractal.h"
#include "Generators/DomainWarp.h"
#include "Generators/DomainWarpSimplex.h"
#include "Generators/DomainWarpFractal.h"
#include "Generators/Modifiers.h"
#include "Generators/Blends.h"

namespace FastNoise
{
    namespace Internal
    {
        // Takes ownership of a freshly constructed node (reference count 0).
        template<typename T>
        SmartNode<T> Adopt( T* node ) { return SmartNode<T>::DynamicCast( ToSmartNode( static_cast<const Generator*>( node ) ) ); }
        FASTNOISE_API bool FeatureSetAllowed( FastSIMD::FeatureSet maxFeatureSet );
    }

    // FastNoise.h:37-43 of the reference; nodes run on the GPU backend.
    template<typename T>
    SmartNode<T> New( FastSIMD::FeatureSet maxFeatureSet )
    {
        static_assert( std::is_base_of<Generator, T>::value, "This function should only be used for FastNoise node classes, for example FastNoise::Simplex" );
        if( !Internal::FeatureSetAllowed( maxFeatureSet ) ) return nullptr;
        return Internal::Adopt( new T );
    }

    // Decodes an encoded node tree (Metadata.cpp:499-506 of the reference); null on failure.
    FASTNOISE_API SmartNode<> NewFromEncodedNodeTree( const char* encodedNodeTreeString, FastSIMD::FeatureSet maxFeatureSet = FastSIMD::FeatureSet::Max );
}
