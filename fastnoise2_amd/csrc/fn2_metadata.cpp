// Metadata registry and wire format of the MI355X backend.
//
// Node ids follow FastNoise2's registration order
// (src/FastNoise/FastSIMD_Build.inl:86-143) and every node's member tables
// follow the construction order of the reference's MetadataT<> constructors
// (base class members first), because encoded node trees and the C ABI
// address members by index. Names, defaults, ranges, groups and enum names are
// the reference's; description texts are this library's own.
//
// Wire format (restated from src/FastNoise/Metadata.cpp:88-308, 310-506 and
// src/FastNoise/Base64.h): a node is `u8 id` (or 0xFF + u16 back-reference),
// then member-lookup bytes {type:3 LSB, index:5}: 0 variable (+i32), 1 node
// lookups (index = count, children follow), 2 hybrid node (+child),
// 3 hybrid constant (+f32), 4 node end (index = run length - 1); only
// non-default values are written. The byte stream is Base64 encoded with
// runs of 'A' compressed to '@' + count.
#include <cstring>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <limits>
#include <mutex>
#include "fn2_internal.h"

using namespace FastNoise;
using namespace FastNoise::Internal;

namespace
{
    struct MetaImpl : Metadata
    {
        Generator* ( *factory )() = nullptr;

        SmartNode<> CreateNode( FastSIMD::FeatureSet level ) const override
        {
            if( !FeatureSetAllowed( level ) ) return nullptr;
            return ToSmartNode( factory() );
        }
    };

    template<typename T> Generator* Make() { return new T; }

    using VU = Metadata::MemberVariable::ValueUnion;

    template<typename T> T* As( Generator* g ) { return dynamic_cast<T*>( g ); }

    // ---------------- member builders ----------------
    template<typename T, typename F>
    void VarF( MetaImpl& m, const char* name, const char* desc, float def, F set, float mn = 0, float mx = 0, int dim = -1, float drag = 0.02f )
    {
        Metadata::MemberVariable v;
        v.name = name; v.description = desc; v.dimensionIdx = dim;
        v.type = Metadata::MemberVariable::EFloat;
        v.valueDefault = def; v.valueMin = mn; v.valueMax = mx; v.valueUiDragSpeed = drag;
        v.setFunc = [set]( Generator* g, VU val ) { if( T* t = As<T>( g ) ) { set( t, val.f ); return true; } return false; };
        m.memberVariables.push_back( v );
    }
    template<typename T, typename F>
    void VarI( MetaImpl& m, const char* name, const char* desc, int def, F set, int mn = 0, int mx = 0 )
    {
        Metadata::MemberVariable v;
        v.name = name; v.description = desc;
        v.type = Metadata::MemberVariable::EInt;
        v.valueDefault = def; v.valueMin = mn; v.valueMax = mx; v.valueUiDragSpeed = 0.2f;
        v.setFunc = [set]( Generator* g, VU val ) { if( T* t = As<T>( g ) ) { set( t, val.i ); return true; } return false; };
        m.memberVariables.push_back( v );
    }
    template<typename T, typename F>
    void VarE( MetaImpl& m, const char* name, const char* desc, int def, F set, std::initializer_list<const char*> names )
    {
        Metadata::MemberVariable v;
        v.name = name; v.description = desc;
        v.type = Metadata::MemberVariable::EEnum;
        v.valueDefault = def;
        v.enumNames.assign( names.begin(), names.end() );
        v.setFunc = [set]( Generator* g, VU val ) { if( T* t = As<T>( g ) ) { set( t, val.i ); return true; } return false; };
        m.memberVariables.push_back( v );
    }
    // node lookup accepting any generator; `set( T*, SmartNodeArg<> )`
    template<typename T, typename F>
    void Lookup( MetaImpl& m, const char* name, const char* desc, F set, int dim = -1 )
    {
        Metadata::MemberNodeLookup l;
        l.name = name; l.description = desc; l.dimensionIdx = dim;
        l.setFunc = [set]( Generator* g, SmartNodeArg<> s ) { if( !s ) return false; if( T* t = As<T>( g ) ) { set( t, s ); return true; } return false; };
        m.memberNodeLookups.push_back( l );
    }
    template<typename T, typename FV, typename FN>
    void Hybrid( MetaImpl& m, const char* name, const char* desc, float def, FV setV, FN setN, int dim = -1, float drag = 0.02f )
    {
        Metadata::MemberHybrid h;
        h.name = name; h.description = desc; h.dimensionIdx = dim;
        h.valueDefault = def; h.valueUiDragSpeed = drag;
        h.setValueFunc = [setV]( Generator* g, float v ) { if( T* t = As<T>( g ) ) { setV( t, v ); return true; } return false; };
        h.setNodeFunc = [setN]( Generator* g, SmartNodeArg<> s ) { if( !s ) return false; if( T* t = As<T>( g ) ) { setN( t, s ); return true; } return false; };
        m.memberHybrids.push_back( h );
    }

    // ---------------- mixin member blocks (reference construction order) ----------------
    template<typename T> void Scalable( MetaImpl& m )
    {
        VarF<T>( m, "Feature Scale", "Size of one noise feature in world units (1 / frequency)", 100.0f, []( T* t, float v ) { t->SetScale( v ); }, 0.f, 0.f, -1, 0.25f );
    }
    template<typename T> void SeededM( MetaImpl& m )
    {
        VarI<T>( m, "Seed Offset", "Added to the seed of this node only; children see the unmodified seed", 0, []( T* t, int v ) { t->SetSeedOffset( v ); } );
    }
    template<typename T> void Range( MetaImpl& m )
    {
        VarF<T>( m, "Output Min", "Lower end of the output range", -1.0f, []( T* t, float v ) { t->SetOutputMin( v ); } );
        VarF<T>( m, "Output Max", "Upper end of the output range", 1.0f, []( T* t, float v ) { t->SetOutputMax( v ); } );
    }
    template<typename T> void CellularM( MetaImpl& m )
    {
        m.groups.push_back( "Coherent Noise" );
        VarE<T>( m, "Distance Function", "Metric used for the distance to cell points", 1, []( T* t, int v ) { t->SetDistanceFunction( (DistanceFunction)v ); },
            { "Euclidean", "Euclidean Squared", "Manhattan", "Hybrid", "Max Axis", "Minkowski" } );
        Hybrid<T>( m, "Minkowski P", "Exponent of the Minkowski metric", 1.5f, []( T* t, float v ) { t->SetMinkowskiP( v ); }, []( T* t, SmartNodeArg<> s ) { t->SetMinkowskiP( s ); } );
        Hybrid<T>( m, "Grid Jitter", "How far cell points move from the cell centre", 1.0f, []( T* t, float v ) { t->SetGridJitter( v ); }, []( T* t, SmartNodeArg<> s ) { t->SetGridJitter( s ); } );
        Hybrid<T>( m, "Size Jitter", "Random per-cell scaling of distances", 0.0f, []( T* t, float v ) { t->SetSizeJitter( v ); }, []( T* t, SmartNodeArg<> s ) { t->SetSizeJitter( s ); } );
    }
    template<typename T, typename SRC = Generator> void FractalM( MetaImpl& m, const char* srcName, bool addGroup )
    {
        if( addGroup ) m.groups.push_back( "Fractal" );
        Metadata::MemberNodeLookup l;
        l.name = srcName; l.description = "Node evaluated once per octave";
        l.setFunc = []( Generator* g, SmartNodeArg<> s ) {
            SmartNode<const SRC> typed = SmartNode<const SRC>::DynamicCast( s );
            if( !typed ) return false;
            if( T* t = As<T>( g ) ) { t->SetSource( typed ); return true; }
            return false;
        };
        m.memberNodeLookups.push_back( l );
        Hybrid<T>( m, "Gain", "Amplitude multiplier per octave", 0.5f, []( T* t, float v ) { t->SetGain( v ); }, []( T* t, SmartNodeArg<> s ) { t->SetGain( s ); } );
        Hybrid<T>( m, "Weighted Strength", "How much an octave's value scales the next octave's amplitude", 0.0f, []( T* t, float v ) { t->SetWeightedStrength( v ); }, []( T* t, SmartNodeArg<> s ) { t->SetWeightedStrength( s ); } );
        VarI<T>( m, "Octaves", "Number of octaves", 3, []( T* t, int v ) { t->SetOctaveCount( v ); }, 2, 16 );
        VarF<T>( m, "Lacunarity", "Frequency multiplier per octave", 2.0f, []( T* t, float v ) { t->SetLacunarity( v ); } );
    }
    template<typename T> void DomainWarpM( MetaImpl& m )
    {
        Scalable<T>( m );
        SeededM<T>( m );
        m.groups.push_back( "Domain Warp" );
        Lookup<T>( m, "Source", "Node evaluated at the warped position", []( T* t, SmartNodeArg<> s ) { t->SetSource( s ); } );
        Hybrid<T>( m, "Warp Amplitude", "Largest distance a position can move", 50.0f, []( T* t, float v ) { t->SetWarpAmplitude( v ); }, []( T* t, SmartNodeArg<> s ) { t->SetWarpAmplitude( s ); }, -1, 0.1f );
        for( int d = 0; d < 4; d++ )
            VarF<T>( m, "Amplitude Scaling", "Per-axis multiplier of the warp amplitude", 1.0f, [d]( T* t, float v ) { t->SetAmplitudeScalingIdx( d, v ); }, 0, 0, d );
    }
    template<typename T> void OpSourceLHS( MetaImpl& m, const char* group )
    {
        m.groups.push_back( group );
        Lookup<T>( m, "LHS", "Left operand", []( T* t, SmartNodeArg<> s ) { t->SetLHS( s ); } );
        Hybrid<T>( m, "RHS", "Right operand", 0.0f, []( T* t, float v ) { t->SetRHS( v ); }, []( T* t, SmartNodeArg<> s ) { t->SetRHS( s ); } );
    }
    template<typename T> void OpHybridLHS( MetaImpl& m, const char* group )
    {
        m.groups.push_back( group );
        Hybrid<T>( m, "LHS", "Left operand", 0.0f, []( T* t, float v ) { t->SetLHS( v ); }, []( T* t, SmartNodeArg<> s ) { t->SetLHS( s ); } );
        Hybrid<T>( m, "RHS", "Right operand", 0.0f, []( T* t, float v ) { t->SetRHS( v ); }, []( T* t, SmartNodeArg<> s ) { t->SetRHS( s ); } );
    }
    template<typename T> void SourceOnly( MetaImpl& m, const char* group )
    {
        m.groups.push_back( group );
        Lookup<T>( m, "Source", "Input node", []( T* t, SmartNodeArg<> s ) { t->SetSource( s ); } );
    }

    // ---------------- the registry ----------------
    struct Registry
    {
        std::vector<MetaImpl*> all;
        std::vector<const Metadata*> view;
        std::unordered_map<std::string, int> byName;

        template<typename T>
        MetaImpl& Add( const char* name, const char* desc )
        {
            auto* m = new MetaImpl;
            m->id = (Metadata::node_id)all.size();
            m->name = name;
            m->description = desc;
            m->factory = &Make<T>;
            all.push_back( m );
            view.push_back( m );
            byName[name] = m->id;
            return *m;
        }

        Registry()
        {
            // Order is the reference's registration order (FastSIMD_Build.inl:86-143).
            {
                auto& m = Add<Constant>( "Constant", "Outputs a constant value" );
                m.groups.push_back( "Basic Generators" );
                VarF<Constant>( m, "Value", "Constant output", 1.0f, []( Constant* t, float v ) { t->SetValue( v ); } );
            }
            {
                auto& m = Add<White>( "White", "White noise: a hash of the exact input position" );
                SeededM<White>( m ); Range<White>( m );
                m.groups.push_back( "Basic Generators" );
            }
            {
                auto& m = Add<Checkerboard>( "Checkerboard", "Alternating cells of Feature Scale size" );
                Scalable<Checkerboard>( m ); Range<Checkerboard>( m );
                m.groups.push_back( "Basic Generators" );
            }
            {
                auto& m = Add<SineWave>( "SineWave", "Product of per-axis sine waves" );
                Scalable<SineWave>( m ); Range<SineWave>( m );
                m.groups.push_back( "Basic Generators" );
            }
            {
                auto& m = Add<Gradient>( "Gradient", "Sum over axes of (position + offset) * multiplier" );
                m.groups.push_back( "Basic Generators" );
                for( int d = 0; d < 4; d++ )
                    VarF<Gradient>( m, "Multiplier", "Per-axis multiplier", 0.0f, [d]( Gradient* t, float v ) { t->SetMultiplierIdx( d, v ); }, 0, 0, d, 0.001f );
                for( int d = 0; d < 4; d++ )
                    Hybrid<Gradient>( m, "Offset", "Per-axis offset", 0.0f, [d]( Gradient* t, float v ) { t->SetOffsetIdx( d, v ); }, [d]( Gradient* t, SmartNodeArg<> s ) { t->SetOffsetIdx( d, s ); }, d, 0.25f );
            }
            {
                auto& m = Add<DistanceToPoint>( "DistanceToPoint", "Distance from the input position to a point" );
                m.groups.push_back( "Basic Generators" );
                VarE<DistanceToPoint>( m, "Distance Function", "Distance metric", 0, []( DistanceToPoint* t, int v ) { t->SetDistanceFunction( (DistanceFunction)v ); },
                    { "Euclidean", "Euclidean Squared", "Manhattan", "Hybrid", "Max Axis", "Minkowski" } );
                for( int d = 0; d < 4; d++ )
                    Hybrid<DistanceToPoint>( m, "Point", "Point coordinate", 0.0f, [d]( DistanceToPoint* t, float v ) { t->SetPointIdx( d, v ); }, [d]( DistanceToPoint* t, SmartNodeArg<> s ) { t->SetPointIdx( d, s ); }, d );
                Hybrid<DistanceToPoint>( m, "Minkowski P", "Exponent of the Minkowski metric", 1.5f, []( DistanceToPoint* t, float v ) { t->SetMinkowskiP( v ); }, []( DistanceToPoint* t, SmartNodeArg<> s ) { t->SetMinkowskiP( s ); } );
            }
            auto lattice = [this]( auto* tag, const char* name, const char* desc ) {
                using T = std::remove_pointer_t<decltype( tag )>;
                auto& m = Add<T>( name, desc );
                Scalable<T>( m ); SeededM<T>( m ); Range<T>( m );
                m.groups.push_back( "Coherent Noise" );
            };
            lattice( (Simplex*)nullptr, "Simplex", "Simplex-lattice gradient noise" );
            lattice( (SuperSimplex*)nullptr, "SuperSimplex", "Smooth simplex-lattice gradient noise (OpenSimplex2S)" );
            lattice( (Perlin*)nullptr, "Perlin", "Square-lattice gradient noise" );
            lattice( (Value*)nullptr, "Value", "Interpolated random values on a square lattice" );
            {
                auto& m = Add<CellularValue>( "CellularValue", "Hash value of the Nth closest cell point" );
                Scalable<CellularValue>( m ); SeededM<CellularValue>( m ); Range<CellularValue>( m ); CellularM<CellularValue>( m );
                VarI<CellularValue>( m, "Value Index", "Which closest cell (0 = nearest)", 0, []( CellularValue* t, int v ) { t->SetValueIndex( v ); }, 0, 3 );
            }
            {
                auto& m = Add<CellularDistance>( "CellularDistance", "Distances to the closest cell points, combined by Return Type" );
                Scalable<CellularDistance>( m ); SeededM<CellularDistance>( m ); Range<CellularDistance>( m ); CellularM<CellularDistance>( m );
                VarI<CellularDistance>( m, "Distance Index 0", "Which closest cell for the first distance", 0, []( CellularDistance* t, int v ) { t->SetDistanceIndex0( v ); }, 0, 3 );
                VarI<CellularDistance>( m, "Distance Index 1", "Which closest cell for the second distance", 1, []( CellularDistance* t, int v ) { t->SetDistanceIndex1( v ); }, 0, 3 );
                VarE<CellularDistance>( m, "Return Type", "How the two distances are combined", 0, []( CellularDistance* t, int v ) { t->SetReturnType( (CellularDistance::ReturnType)v ); },
                    { "Index0", "Index0Add1", "Index0Sub1", "Index0Mul1", "Index0Div1" } );
            }
            {
                auto& m = Add<CellularLookup>( "CellularLookup", "Lookup node evaluated at the closest cell point" );
                Scalable<CellularLookup>( m ); SeededM<CellularLookup>( m ); CellularM<CellularLookup>( m );
                Lookup<CellularLookup>( m, "Lookup", "Node evaluated at the cell point", []( CellularLookup* t, SmartNodeArg<> s ) { t->SetLookup( s ); } );
            }
            {
                auto& m = Add<FractalFBm>( "FractalFBm", "Fractional Brownian motion: sum of octaves of the source" );
                FractalM<FractalFBm>( m, "Source", true );
            }
            {
                auto& m = Add<PingPong>( "PingPong", "Folds the scaled source back and forth between 0 and 1" );
                SourceOnly<PingPong>( m, "Modifiers" );
                Hybrid<PingPong>( m, "Ping Pong Strength", "Multiplier applied before folding", 2.0f, []( PingPong* t, float v ) { t->SetPingPongStrength( v ); }, []( PingPong* t, SmartNodeArg<> s ) { t->SetPingPongStrength( s ); } );
            }
            {
                auto& m = Add<FractalRidged>( "FractalRidged", "Ridged fractal: octaves of 1 - 2|source|" );
                FractalM<FractalRidged>( m, "Source", true );
            }
            auto simplexWarp = [this]( auto* tag, const char* name, const char* desc ) {
                using T = std::remove_pointer_t<decltype( tag )>;
                auto& m = Add<T>( name, desc );
                DomainWarpM<T>( m );
                VarE<T>( m, "Vectorization Scheme", "How the noise produces a vector", 0, []( T* t, int v ) { t->SetVectorizationScheme( (VectorizationScheme)v ); },
                    { "Orthogonal Gradient Matrix", "Gradient Outer Product" } );
            };
            simplexWarp( (DomainWarpSimplex*)nullptr, "DomainWarpSimplex", "Warps the position with simplex-lattice vector noise" );
            simplexWarp( (DomainWarpSuperSimplex*)nullptr, "DomainWarpSuperSimplex", "Warps the position with smooth simplex-lattice vector noise" );
            {
                auto& m = Add<DomainWarpGradient>( "DomainWarpGradient", "Warps the position with square-lattice vector noise" );
                DomainWarpM<DomainWarpGradient>( m );
            }
            auto warpFractal = [this]( auto* tag, const char* name, const char* desc ) {
                using T = std::remove_pointer_t<decltype( tag )>;
                auto& m = Add<T>( name, desc );
                FractalM<T, DomainWarp>( m, "Domain Warp Source", false );
                m.groups.push_back( "Domain Warp" );
                m.groups.push_back( "Fractal" );
            };
            warpFractal( (DomainWarpFractalProgressive*)nullptr, "DomainWarpFractalProgressive", "Applies the warp once per octave, each octave warping the previous result" );
            warpFractal( (DomainWarpFractalIndependent*)nullptr, "DomainWarpFractalIndependent", "Applies the warp once per octave from the original position and sums the offsets" );
            {
                auto& m = Add<FastNoise::Add>( "Add", "LHS + RHS" );
                OpSourceLHS<FastNoise::Add>( m, "Operators" );
            }
            {
                auto& m = Add<Subtract>( "Subtract", "LHS - RHS" );
                OpHybridLHS<Subtract>( m, "Operators" );
            }
            {
                auto& m = Add<Multiply>( "Multiply", "LHS * RHS" );
                OpSourceLHS<Multiply>( m, "Operators" );
            }
            {
                auto& m = Add<Divide>( "Divide", "LHS / RHS" );
                OpHybridLHS<Divide>( m, "Operators" );
            }
            {
                auto& m = Add<Abs>( "Abs", "Absolute value of the source" );
                SourceOnly<Abs>( m, "Modifiers" );
            }
            {
                auto& m = Add<FastNoise::Min>( "Min", "Smaller of LHS and RHS" );
                OpSourceLHS<FastNoise::Min>( m, "Blends" );
            }
            {
                auto& m = Add<FastNoise::Max>( "Max", "Larger of LHS and RHS" );
                OpSourceLHS<FastNoise::Max>( m, "Blends" );
            }
            {
                auto& m = Add<MinSmooth>( "MinSmooth", "Smooth minimum of LHS and RHS" );
                OpSourceLHS<MinSmooth>( m, "Blends" );
                Hybrid<MinSmooth>( m, "Smoothness", "Width of the smooth transition", 0.1f, []( MinSmooth* t, float v ) { t->SetSmoothness( v ); }, []( MinSmooth* t, SmartNodeArg<> s ) { t->SetSmoothness( s ); } );
            }
            {
                auto& m = Add<MaxSmooth>( "MaxSmooth", "Smooth maximum of LHS and RHS" );
                OpSourceLHS<MaxSmooth>( m, "Blends" );
                Hybrid<MaxSmooth>( m, "Smoothness", "Width of the smooth transition", 0.1f, []( MaxSmooth* t, float v ) { t->SetSmoothness( v ); }, []( MaxSmooth* t, SmartNodeArg<> s ) { t->SetSmoothness( s ); } );
            }
            {
                auto& m = Add<SignedSquareRoot>( "SignedSquareRoot", "Square root of |source| with the sign of the source" );
                SourceOnly<SignedSquareRoot>( m, "Modifiers" );
            }
            {
                auto& m = Add<PowFloat>( "PowFloat", "Value raised to the power Pow" );
                m.groups.push_back( "Blends" );
                Hybrid<PowFloat>( m, "Value", "Base", 2.0f, []( PowFloat* t, float v ) { t->SetValue( v ); }, []( PowFloat* t, SmartNodeArg<> s ) { t->SetValue( s ); } );
                Hybrid<PowFloat>( m, "Pow", "Exponent", 2.0f, []( PowFloat* t, float v ) { t->SetPow( v ); }, []( PowFloat* t, SmartNodeArg<> s ) { t->SetPow( s ); } );
            }
            {
                auto& m = Add<PowInt>( "PowInt", "Value raised to an integer power" );
                m.groups.push_back( "Blends" );
                Lookup<PowInt>( m, "Value", "Base", []( PowInt* t, SmartNodeArg<> s ) { t->SetValue( s ); } );
                VarI<PowInt>( m, "Pow", "Integer exponent", 2, []( PowInt* t, int v ) { t->SetPow( v ); }, 2, 0 );
            }
            {
                auto& m = Add<DomainScale>( "DomainScale", "Scales the input position" );
                SourceOnly<DomainScale>( m, "Domain Modifiers" );
                VarF<DomainScale>( m, "Scaling", "Position multiplier", 1.0f, []( DomainScale* t, float v ) { t->SetScaling( v ); } );
            }
            {
                auto& m = Add<DomainOffset>( "DomainOffset", "Offsets the input position" );
                SourceOnly<DomainOffset>( m, "Domain Modifiers" );
                for( int d = 0; d < 4; d++ )
                    Hybrid<DomainOffset>( m, "Offset", "Per-axis offset", 0.0f, [d]( DomainOffset* t, float v ) { t->SetOffsetIdx( d, v ); }, [d]( DomainOffset* t, SmartNodeArg<> s ) { t->SetOffsetIdx( d, s ); }, d, 0.25f );
            }
            {
                auto& m = Add<DomainRotate>( "DomainRotate", "Rotates the input position (yaw, pitch, roll in radians)" );
                SourceOnly<DomainRotate>( m, "Domain Modifiers" );
                VarF<DomainRotate>( m, "Yaw", "Rotation about Z", 0.0f, []( DomainRotate* t, float v ) { t->SetYaw( v ); } );
                VarF<DomainRotate>( m, "Pitch", "Rotation about Y", 0.0f, []( DomainRotate* t, float v ) { t->SetPitch( v ); } );
                VarF<DomainRotate>( m, "Roll", "Rotation about X", 0.0f, []( DomainRotate* t, float v ) { t->SetRoll( v ); } );
            }
            {
                auto& m = Add<DomainAxisScale>( "DomainAxisScale", "Scales each axis of the input position" );
                SourceOnly<DomainAxisScale>( m, "Domain Modifiers" );
                for( int d = 0; d < 4; d++ )
                    VarF<DomainAxisScale>( m, "Scaling", "Per-axis multiplier", 1.0f, [d]( DomainAxisScale* t, float v ) { t->SetScalingIdx( d, v ); }, 0, 0, d );
            }
            {
                auto& m = Add<SeedOffset>( "SeedOffset", "Adds to the seed passed to the source" );
                SourceOnly<SeedOffset>( m, "Modifiers" );
                VarI<SeedOffset>( m, "Seed Offset", "Added to the seed", 1, []( SeedOffset* t, int v ) { t->SetOffset( v ); } );
            }
            {
                auto& m = Add<ConvertRGBA8>( "ConvertRGBA8", "Packs the source into a grey RGBA8 value stored in the float's bits" );
                SourceOnly<ConvertRGBA8>( m, "Modifiers" );
                VarF<ConvertRGBA8>( m, "Min", "Source value mapped to 0", -1.0f, []( ConvertRGBA8* t, float v ) { t->SetMin( v ); } );
                VarF<ConvertRGBA8>( m, "Max", "Source value mapped to 255", 1.0f, []( ConvertRGBA8* t, float v ) { t->SetMax( v ); } );
            }
            {
                auto& m = Add<GeneratorCache>( "GeneratorCache", "Pass-through; every point is evaluated once on the device" );
                SourceOnly<GeneratorCache>( m, "Modifiers" );
            }
            {
                auto& m = Add<Fade>( "Fade", "Interpolates from A to B by the Fade input" );
                m.groups.push_back( "Blends" );
                Lookup<Fade>( m, "A", "From", []( Fade* t, SmartNodeArg<> s ) { t->SetA( s ); } );
                Lookup<Fade>( m, "B", "To", []( Fade* t, SmartNodeArg<> s ) { t->SetB( s ); } );
                Hybrid<Fade>( m, "Fade", "Interpolation position", 0.0f, []( Fade* t, float v ) { t->SetFade( v ); }, []( Fade* t, SmartNodeArg<> s ) { t->SetFade( s ); } );
                Hybrid<Fade>( m, "Fade Min", "Fade value giving A", -1.0f, []( Fade* t, float v ) { t->SetFadeMin( v ); }, []( Fade* t, SmartNodeArg<> s ) { t->SetFadeMin( s ); } );
                Hybrid<Fade>( m, "Fade Max", "Fade value giving B", 1.0f, []( Fade* t, float v ) { t->SetFadeMax( v ); }, []( Fade* t, SmartNodeArg<> s ) { t->SetFadeMax( s ); } );
                VarE<Fade>( m, "Interpolation", "Easing curve", 0, []( Fade* t, int v ) { t->SetInterpolation( (Fade::Interpolation)v ); }, { "Linear", "Hermite", "Quintic" } );
            }
            {
                auto& m = Add<Remap>( "Remap", "Linearly maps the source from one range to another" );
                SourceOnly<Remap>( m, "Modifiers" );
                Hybrid<Remap>( m, "From Min", "Input range start", -1.0f, []( Remap* t, float v ) { t->SetFromMin( v ); }, []( Remap* t, SmartNodeArg<> s ) { t->SetFromMin( s ); } );
                Hybrid<Remap>( m, "From Max", "Input range end", 1.0f, []( Remap* t, float v ) { t->SetFromMax( v ); }, []( Remap* t, SmartNodeArg<> s ) { t->SetFromMax( s ); } );
                Hybrid<Remap>( m, "To Min", "Output range start", 0.0f, []( Remap* t, float v ) { t->SetToMin( v ); }, []( Remap* t, SmartNodeArg<> s ) { t->SetToMin( s ); } );
                Hybrid<Remap>( m, "To Max", "Output range end", 1.0f, []( Remap* t, float v ) { t->SetToMax( v ); }, []( Remap* t, SmartNodeArg<> s ) { t->SetToMax( s ); } );
                VarE<Remap>( m, "Clamp Output", "Clamp to the output range", 0, []( Remap* t, int v ) { t->SetClampOutput( (Boolean)v ); }, { "False", "True" } );
            }
            {
                auto& m = Add<Terrace>( "Terrace", "Quantises the source into steps" );
                SourceOnly<Terrace>( m, "Modifiers" );
                VarF<Terrace>( m, "Step Count", "Steps per unit", 1.0f, []( Terrace* t, float v ) { t->SetStepCount( v ); } );
                Hybrid<Terrace>( m, "Smoothness", "Softening of the step edges", 0.0f, []( Terrace* t, float v ) { t->SetSmoothness( v ); }, []( Terrace* t, SmartNodeArg<> s ) { t->SetSmoothness( s ); } );
            }
            {
                auto& m = Add<AddDimension>( "AddDimension", "Evaluates the source with one more dimension" );
                SourceOnly<AddDimension>( m, "Domain Modifiers" );
                Hybrid<AddDimension>( m, "New Dimension Position", "Coordinate of the added axis", 0.0f, []( AddDimension* t, float v ) { t->SetNewDimensionPosition( v ); }, []( AddDimension* t, SmartNodeArg<> s ) { t->SetNewDimensionPosition( s ); } );
            }
            {
                auto& m = Add<RemoveDimension>( "RemoveDimension", "Evaluates the source without one axis" );
                SourceOnly<RemoveDimension>( m, "Domain Modifiers" );
                VarE<RemoveDimension>( m, "Remove Dimension", "Axis to drop", 1, []( RemoveDimension* t, int v ) { t->SetRemoveDimension( (Dim)v ); }, { "X", "Y", "Z", "W" } );
            }
            {
                auto& m = Add<Modulus>( "Modulus", "Floating-point remainder of LHS / RHS" );
                OpHybridLHS<Modulus>( m, "Operators" );
            }
            {
                auto& m = Add<DomainRotatePlane>( "DomainRotatePlane", "Rotation that improves the look of planar slices" );
                SourceOnly<DomainRotatePlane>( m, "Domain Modifiers" );
                VarE<DomainRotatePlane>( m, "Rotation Type", "Plane to improve", 0, []( DomainRotatePlane* t, int v ) { t->SetRotationType( (PlaneRotationType)v ); }, { "Improve XY Planes", "Improve XZ Planes" } );
            }
        }
    };

    Registry& Reg()
    {
        static Registry r;
        return r;
    }

    // ---------------- Base64 with 'A' run-length ----------------
    const char kB64[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

    int B64Val( unsigned char c )
    {
        if( c >= 'A' && c <= 'Z' ) return c - 'A';
        if( c >= 'a' && c <= 'z' ) return c - 'a' + 26;
        if( c >= '0' && c <= '9' ) return c - '0' + 52;
        if( c == '+' ) return 62;
        if( c == '/' ) return 63;
        if( c == '=' ) return 0;
        return 64;
    }

    bool B64Decode( const char* in, std::vector<uint8_t>& out )
    {
        if( !in ) return false;
        std::string plain;
        size_t raw = 0;
        for( ; in[raw]; )
        {
            if( in[raw] == '@' )
            {
                unsigned char c = (unsigned char)in[raw + 1];
                int extra = B64Val( c );
                if( c == 0 || extra == 64 ) return false;
                plain.append( (size_t)extra + 3, 'A' );
                raw += 2;
            }
            else plain += in[raw++];
        }
        if( plain.empty() || plain.size() % 4 ) return false;
        size_t n = plain.size() / 4 * 3;
        if( in[raw - 1] == '=' ) { n--; if( raw > 1 && in[raw - 2] == '=' ) n--; }
        out.resize( n );
        size_t j = 0;
        for( size_t i = 0; i < plain.size(); i += 4 )
        {
            uint32_t t = ( (uint32_t)B64Val( plain[i] ) << 18 ) | ( (uint32_t)B64Val( plain[i + 1] ) << 12 ) |
                         ( (uint32_t)B64Val( plain[i + 2] ) << 6 ) | (uint32_t)B64Val( plain[i + 3] );
            for( int b = 2; b >= 0 && j < n; b-- ) out[j++] = ( t >> ( 8 * b ) ) & 0xFF;
        }
        return true;
    }

    std::string B64Encode( const std::vector<uint8_t>& data )
    {
        std::string plain;
        size_t i = 0;
        for( ; i + 2 < data.size(); i += 3 )
        {
            uint32_t t = ( data[i] << 16 ) | ( data[i + 1] << 8 ) | data[i + 2];
            for( int b = 3; b >= 0; b-- ) plain += kB64[( t >> ( 6 * b ) ) & 63];
        }
        if( i < data.size() )
        {
            uint32_t t = data[i] << 16;
            if( i + 1 < data.size() ) t |= data[i + 1] << 8;
            plain += kB64[( t >> 18 ) & 63];
            plain += kB64[( t >> 12 ) & 63];
            plain += i + 1 < data.size() ? kB64[( t >> 6 ) & 63] : '=';
            plain += '=';
        }
        // runs of 3..66 'A' become '@' + count-3
        std::string out;
        for( size_t k = 0; k < plain.size(); )
        {
            size_t run = 0;
            while( k + run < plain.size() && plain[k + run] == 'A' && run < 66 ) run++;
            if( run >= 3 ) { out += '@'; out += kB64[run - 3]; k += run; }
            else if( run > 0 ) { out.append( run, 'A' ); k += run; }
            else out += plain[k++];
        }
        return out;
    }

    // ---------------- node stream ----------------
    struct InStream
    {
        std::vector<uint8_t> b;
        size_t pos = 0;
        int pendingEnds = 0;

        template<typename T> bool Get( T& v )
        {
            if( pendingEnds > 0 || pos + sizeof( T ) > b.size() ) return false;
            std::memcpy( &v, b.data() + pos, sizeof( T ) );
            pos += sizeof( T );
            return true;
        }
        bool Member( int& type, int& index )
        {
            if( pendingEnds > 0 ) { pendingEnds--; type = 4; index = 0; return true; }
            if( pos >= b.size() ) return false;
            uint8_t m = b[pos++];
            type = m & 7;
            index = m >> 3;
            if( type == 4 ) pendingEnds = index;
            return true;
        }
    };

    SmartNode<> DecodeNode( InStream& s, std::vector<SmartNode<>>& refs, FastSIMD::FeatureSet level )
    {
        uint8_t id;
        if( !s.Get( id ) ) return nullptr;
        if( id == 0xFF )
        {
            uint16_t ref;
            if( !s.Get( ref ) || ref >= refs.size() ) return nullptr;
            return refs[ref];
        }
        const Metadata* meta = Metadata::GetFromId( id );
        if( !meta ) return nullptr;
        SmartNode<> node = meta->CreateNode( level );
        if( !node ) return nullptr;
        int type, index;
        if( !s.Member( type, index ) ) return nullptr;
        while( type == 0 )
        {
            VU v;
            if( !s.Get( v.i ) ) return nullptr;
            if( (size_t)index < meta->memberVariables.size() ) meta->memberVariables[index].setFunc( node.get(), v );
            if( !s.Member( type, index ) ) return nullptr;
        }
        if( type == 1 )
        {
            size_t i = 0;
            for( ; i < std::min<size_t>( index, meta->memberNodeLookups.size() ); i++ )
            {
                SmartNode<> c = DecodeNode( s, refs, level );
                if( !c || !meta->memberNodeLookups[i].setFunc( node.get(), c ) ) return nullptr;
            }
            for( ; i < (size_t)index; i++ )
                if( !DecodeNode( s, refs, level ) ) return nullptr;
            for( ; i < meta->memberNodeLookups.size(); i++ )
                if( !meta->memberNodeLookups[i].setFunc( node.get(), New<Constant>( level ) ) ) return nullptr;
            if( !s.Member( type, index ) ) return nullptr;
        }
        while( type != 4 )
        {
            if( type == 2 )
            {
                SmartNode<> c = DecodeNode( s, refs, level );
                if( (size_t)index < meta->memberHybrids.size() )
                    if( !c || !meta->memberHybrids[index].setNodeFunc( node.get(), c ) ) return nullptr;
            }
            else if( type == 3 )
            {
                float v;
                if( !s.Get( v ) ) return nullptr;
                if( (size_t)index < meta->memberHybrids.size() ) meta->memberHybrids[index].setValueFunc( node.get(), v );
            }
            if( !s.Member( type, index ) ) return nullptr;
        }
        refs.push_back( node );
        return node;
    }

    struct OutStream
    {
        std::vector<uint8_t> b;
        int pendingEnds = 0;

        void Flush()
        {
            if( pendingEnds > 0 )
            {
                b.push_back( (uint8_t)( 4 | ( ( pendingEnds - 1 ) << 3 ) ) );
                pendingEnds = 0;
            }
        }
        template<typename T> void Put( T v )
        {
            Flush();
            for( size_t i = 0; i < sizeof( T ); i++ ) b.push_back( (uint8_t)( (uint64_t)v >> ( 8 * i ) ) );
        }
        void Member( int type, int index ) { Put<uint8_t>( (uint8_t)( type | ( index << 3 ) ) ); }
        void End()
        {
            if( ++pendingEnds >= 32 ) Flush();
        }
    };

    bool EncodeNodeData( NodeData* nd, bool fixUp, OutStream& s, std::unordered_map<const NodeData*, uint16_t>& refs, std::unordered_set<const NodeData*> deps )
    {
        const Metadata* meta = nd->metadata;
        if( !meta || nd->variables.size() != meta->memberVariables.size() || nd->nodeLookups.size() != meta->memberNodeLookups.size() ||
            nd->hybrids.size() != meta->memberHybrids.size() )
            return false;
        if( fixUp )
        {
            deps.insert( nd );
            for( auto& n : nd->nodeLookups ) if( deps.count( n ) ) n = nullptr;
            for( auto& h : nd->hybrids ) if( deps.count( h.first ) ) h.first = nullptr;
        }
        auto it = refs.find( nd );
        if( it != refs.end() )
        {
            s.Put<uint8_t>( 0xFF );
            s.Put<uint16_t>( it->second );
            return true;
        }
        s.Put<uint8_t>( meta->id );
        for( size_t i = 0; i < meta->memberVariables.size(); i++ )
        {
            if( nd->variables[i].i != meta->memberVariables[i].valueDefault.i )
            {
                s.Member( 0, (int)i );
                s.Put<int32_t>( nd->variables[i].i );
            }
        }
        if( meta->memberNodeLookups.size() ) s.Member( 1, (int)meta->memberNodeLookups.size() );
        for( size_t i = 0; i < meta->memberNodeLookups.size(); i++ )
        {
            if( fixUp && nd->nodeLookups[i] )
            {
                SmartNode<> test = meta->CreateNode();
                SmartNode<> src = nd->nodeLookups[i]->metadata->CreateNode();
                if( !meta->memberNodeLookups[i].setFunc( test.get(), src ) ) { nd->nodeLookups[i] = nullptr; return false; }
            }
            if( !nd->nodeLookups[i] || !EncodeNodeData( nd->nodeLookups[i], fixUp, s, refs, deps ) ) return false;
        }
        for( size_t i = 0; i < meta->memberHybrids.size(); i++ )
        {
            if( !nd->hybrids[i].first )
            {
                if( nd->hybrids[i].second != meta->memberHybrids[i].valueDefault )
                {
                    s.Member( 3, (int)i );
                    VU v = nd->hybrids[i].second;
                    s.Put<int32_t>( v.i );
                }
            }
            else
            {
                if( fixUp )
                {
                    SmartNode<> test = meta->CreateNode();
                    SmartNode<> src = nd->hybrids[i].first->metadata->CreateNode();
                    if( !meta->memberHybrids[i].setNodeFunc( test.get(), src ) ) { nd->hybrids[i].first = nullptr; return false; }
                }
                s.Member( 2, (int)i );
                if( !EncodeNodeData( nd->hybrids[i].first, fixUp, s, refs, deps ) ) return false;
            }
        }
        s.End();
        refs.emplace( nd, (uint16_t)refs.size() );
        return true;
    }

    NodeData* DecodeNodeData( InStream& s, std::vector<std::unique_ptr<NodeData>>& out )
    {
        uint8_t id;
        if( !s.Get( id ) ) return nullptr;
        if( id == 0xFF )
        {
            uint16_t ref;
            if( !s.Get( ref ) || ref >= out.size() ) return nullptr;
            return out[ref].get();
        }
        const Metadata* meta = Metadata::GetFromId( id );
        if( !meta ) return nullptr;
        std::unique_ptr<NodeData> nd( new NodeData( meta ) );
        int type, index;
        if( !s.Member( type, index ) ) return nullptr;
        while( type == 0 )
        {
            VU v;
            if( !s.Get( v.i ) ) return nullptr;
            if( (size_t)index < meta->memberVariables.size() ) nd->variables[index] = v;
            if( !s.Member( type, index ) ) return nullptr;
        }
        if( type == 1 )
        {
            size_t i = 0;
            for( ; i < std::min<size_t>( index, meta->memberNodeLookups.size() ); i++ ) nd->nodeLookups[i] = DecodeNodeData( s, out );
            for( ; i < (size_t)index; i++ ) DecodeNodeData( s, out );
            if( !s.Member( type, index ) ) return nullptr;
        }
        while( type != 4 )
        {
            if( type == 2 )
            {
                NodeData* c = DecodeNodeData( s, out );
                if( (size_t)index < meta->memberHybrids.size() ) nd->hybrids[index].first = c;
            }
            else if( type == 3 )
            {
                float v;
                if( !s.Get( v ) ) return nullptr;
                if( (size_t)index < meta->memberHybrids.size() ) nd->hybrids[index].second = v;
            }
            if( !s.Member( type, index ) ) return nullptr;
        }
        out.emplace_back( std::move( nd ) );
        return out.back().get();
    }
}

namespace FastNoise
{
    const std::vector<const Metadata*>& Metadata::GetAll() { return Reg().view; }

    const Metadata* Metadata::GetFromId( node_id nodeId )
    {
        auto& v = Reg().view;
        return nodeId < v.size() ? v[nodeId] : nullptr;
    }

    std::string Metadata::SerialiseNodeData( NodeData* nodeData, bool fixUp )
    {
        OutStream s;
        std::unordered_map<const NodeData*, uint16_t> refs;
        if( !nodeData || !EncodeNodeData( nodeData, fixUp, s, refs, {} ) ) return "";
        s.Flush();
        return B64Encode( s.b );
    }

    NodeData* Metadata::DeserialiseNodeData( const char* serialised, std::vector<std::unique_ptr<NodeData>>& nodeDataOut )
    {
        InStream s;
        if( !B64Decode( serialised, s.b ) ) return nullptr;
        return DecodeNodeData( s, nodeDataOut );
    }

    std::string Metadata::FormatMetadataNodeName( const Metadata* metadata, bool removeGroups )
    {
        std::string str;
        if( metadata->formattedName ) str = metadata->formattedName;
        else
        {
            str = metadata->name;
            for( size_t i = 1; i < str.size(); i++ )
                if( ( isdigit( (unsigned char)str[i] ) || isupper( (unsigned char)str[i] ) ) && islower( (unsigned char)str[i - 1] ) )
                    str.insert( i++, 1, ' ' );
        }
        if( removeGroups )
        {
            for( const char* group : metadata->groups )
            {
                size_t p = str.find( group );
                if( p != std::string::npos ) str.erase( p, std::strlen( group ) + 1 );
            }
        }
        if( str.empty() ) return metadata->name;
        return str;
    }

    std::string Metadata::FormatMetadataMemberName( const Member& member )
    {
        std::string str = member.name;
        if( member.dimensionIdx >= 0 )
        {
            str.insert( str.begin(), ' ' );
            str.insert( 0, kDim_Strings[member.dimensionIdx] );
        }
        return str;
    }

    SmartNode<> NewFromEncodedNodeTree( const char* encoded, FastSIMD::FeatureSet level )
    {
        InStream s;
        if( !B64Decode( encoded, s.b ) ) return nullptr;
        std::vector<SmartNode<>> refs;
        return DecodeNode( s, refs, level );
    }

    namespace Impl
    {
#define FN2_META_GET( T ) template<> const Metadata& GetMetadata<T>() { return *Metadata::GetFromId( (Metadata::node_id)Internal::k##T ); }
        FN2_META_GET( Constant ) FN2_META_GET( White ) FN2_META_GET( Checkerboard ) FN2_META_GET( SineWave ) FN2_META_GET( Gradient )
        FN2_META_GET( DistanceToPoint ) FN2_META_GET( Simplex ) FN2_META_GET( SuperSimplex ) FN2_META_GET( Perlin ) FN2_META_GET( Value )
        FN2_META_GET( CellularValue ) FN2_META_GET( CellularDistance ) FN2_META_GET( CellularLookup ) FN2_META_GET( FractalFBm )
        FN2_META_GET( PingPong ) FN2_META_GET( FractalRidged ) FN2_META_GET( DomainWarpSimplex ) FN2_META_GET( DomainWarpSuperSimplex )
        FN2_META_GET( DomainWarpGradient ) FN2_META_GET( DomainWarpFractalProgressive ) FN2_META_GET( DomainWarpFractalIndependent )
        FN2_META_GET( Add ) FN2_META_GET( Subtract ) FN2_META_GET( Multiply ) FN2_META_GET( Divide ) FN2_META_GET( Abs ) FN2_META_GET( Min )
        FN2_META_GET( Max ) FN2_META_GET( MinSmooth ) FN2_META_GET( MaxSmooth ) FN2_META_GET( SignedSquareRoot ) FN2_META_GET( PowFloat )
        FN2_META_GET( PowInt ) FN2_META_GET( DomainScale ) FN2_META_GET( DomainOffset ) FN2_META_GET( DomainRotate ) FN2_META_GET( DomainAxisScale )
        FN2_META_GET( SeedOffset ) FN2_META_GET( ConvertRGBA8 ) FN2_META_GET( GeneratorCache ) FN2_META_GET( Fade ) FN2_META_GET( Remap )
        FN2_META_GET( Terrace ) FN2_META_GET( AddDimension ) FN2_META_GET( RemoveDimension ) FN2_META_GET( Modulus ) FN2_META_GET( DomainRotatePlane )
#undef FN2_META_GET
    }

#define FN2_NODE_META( T ) const Metadata& T::GetMetadata() const { return Metadata::Get<T>(); }
    FN2_NODE_META( Constant ) FN2_NODE_META( White ) FN2_NODE_META( Checkerboard ) FN2_NODE_META( SineWave ) FN2_NODE_META( Gradient )
    FN2_NODE_META( DistanceToPoint ) FN2_NODE_META( Simplex ) FN2_NODE_META( SuperSimplex ) FN2_NODE_META( Perlin ) FN2_NODE_META( Value )
    FN2_NODE_META( CellularValue ) FN2_NODE_META( CellularDistance ) FN2_NODE_META( CellularLookup ) FN2_NODE_META( FractalFBm )
    FN2_NODE_META( PingPong ) FN2_NODE_META( FractalRidged ) FN2_NODE_META( DomainWarpSimplex ) FN2_NODE_META( DomainWarpSuperSimplex )
    FN2_NODE_META( DomainWarpGradient ) FN2_NODE_META( DomainWarpFractalProgressive ) FN2_NODE_META( DomainWarpFractalIndependent )
    FN2_NODE_META( Add ) FN2_NODE_META( Subtract ) FN2_NODE_META( Multiply ) FN2_NODE_META( Divide ) FN2_NODE_META( Abs ) FN2_NODE_META( Min )
    FN2_NODE_META( Max ) FN2_NODE_META( MinSmooth ) FN2_NODE_META( MaxSmooth ) FN2_NODE_META( SignedSquareRoot ) FN2_NODE_META( PowFloat )
    FN2_NODE_META( PowInt ) FN2_NODE_META( DomainScale ) FN2_NODE_META( DomainOffset ) FN2_NODE_META( DomainRotate ) FN2_NODE_META( DomainAxisScale )
    FN2_NODE_META( SeedOffset ) FN2_NODE_META( ConvertRGBA8 ) FN2_NODE_META( GeneratorCache ) FN2_NODE_META( Fade ) FN2_NODE_META( Remap )
    FN2_NODE_META( Terrace ) FN2_NODE_META( AddDimension ) FN2_NODE_META( RemoveDimension ) FN2_NODE_META( Modulus ) FN2_NODE_META( DomainRotatePlane )
#undef FN2_NODE_META
}
