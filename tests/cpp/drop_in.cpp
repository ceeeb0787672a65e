// A C++ caller of the drop-in API, compiled against include/FastNoise and
// linked with libfastnoise2_amd.so the way a FastNoise2 user relinks: node
// construction through New<T> for every node class, typed setters,
// SmartNode ownership, encoded trees, and the batch entry points into host
// std::vectors (FastNoise.h:37-65 and Generator.h:210-429 of the reference,
// in the shape of its tests/FastNoiseCpp11Include.cpp).
//
//   drop_in <output dir> <encoded FBm(Simplex) tree>
//
// Checks that need no GPU run first and stop the program with exit status 1
// and a message when they fail. The generation calls then write their
// outputs (raw float32) and ranges into the output directory; without a GPU
// those calls fill their outputs with NaN, which the caller checks.
#include <FastNoise/FastNoise.h>
#include <FastNoise/Metadata.h>

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#define CHECK( c ) \
    do { if( !( c ) ) { std::fprintf( stderr, "check failed at line %d: %s\n", __LINE__, #c ); return 1; } } while( 0 )

namespace
{
    std::string gDir;

    bool Save( const char* name, const std::vector<float>& v, FastNoise::OutputMinMax mm )
    {
        std::string path = gDir + "/" + name + ".f32";
        FILE* f = std::fopen( path.c_str(), "wb" );
        if( !f ) return false;
        std::fwrite( v.data(), sizeof( float ), v.size(), f );
        std::fclose( f );
        path = gDir + "/" + name + ".range";
        f = std::fopen( path.c_str(), "wb" );
        if( !f ) return false;
        float r[2] = { mm.min, mm.max };
        std::fwrite( r, sizeof( float ), 2, f );
        std::fclose( f );
        return true;
    }

    // New<T> with the node id of the reference's registration order
    // (FastSIMD_Build.inl:86-143)
    template<typename T>
    int Make( int expectedId )
    {
        FastNoise::SmartNode<T> n = FastNoise::New<T>();
        CHECK( n );
        CHECK( n->GetMetadata().id == expectedId );
        CHECK( n.use_count() == 1 );
        FastNoise::SmartNode<> base = n;
        CHECK( n.use_count() == 2 && base.use_count() == 2 );
        CHECK( FastNoise::SmartNode<T>::DynamicCast( base ) == n );
        return 0;
    }
}

int main( int argc, char** argv )
{
    if( argc < 3 )
    {
        std::fprintf( stderr, "usage: drop_in <output dir> <encoded tree>\n" );
        return 2;
    }
    gDir = argv[1];
    using namespace FastNoise;

    // ---- every node class through New<T>
    int id = 0, bad = 0;
    bad |= Make<Constant>( id++ ); bad |= Make<White>( id++ ); bad |= Make<Checkerboard>( id++ );
    bad |= Make<SineWave>( id++ ); bad |= Make<Gradient>( id++ ); bad |= Make<DistanceToPoint>( id++ );
    bad |= Make<Simplex>( id++ ); bad |= Make<SuperSimplex>( id++ ); bad |= Make<Perlin>( id++ ); bad |= Make<Value>( id++ );
    bad |= Make<CellularValue>( id++ ); bad |= Make<CellularDistance>( id++ ); bad |= Make<CellularLookup>( id++ );
    bad |= Make<FractalFBm>( id++ ); bad |= Make<PingPong>( id++ ); bad |= Make<FractalRidged>( id++ );
    bad |= Make<DomainWarpSimplex>( id++ ); bad |= Make<DomainWarpSuperSimplex>( id++ ); bad |= Make<DomainWarpGradient>( id++ );
    bad |= Make<DomainWarpFractalProgressive>( id++ ); bad |= Make<DomainWarpFractalIndependent>( id++ );
    bad |= Make<Add>( id++ ); bad |= Make<Subtract>( id++ ); bad |= Make<Multiply>( id++ ); bad |= Make<Divide>( id++ );
    bad |= Make<Abs>( id++ ); bad |= Make<Min>( id++ ); bad |= Make<Max>( id++ ); bad |= Make<MinSmooth>( id++ );
    bad |= Make<MaxSmooth>( id++ ); bad |= Make<SignedSquareRoot>( id++ ); bad |= Make<PowFloat>( id++ ); bad |= Make<PowInt>( id++ );
    bad |= Make<DomainScale>( id++ ); bad |= Make<DomainOffset>( id++ ); bad |= Make<DomainRotate>( id++ );
    bad |= Make<DomainAxisScale>( id++ ); bad |= Make<SeedOffset>( id++ ); bad |= Make<ConvertRGBA8>( id++ );
    bad |= Make<GeneratorCache>( id++ ); bad |= Make<Fade>( id++ ); bad |= Make<Remap>( id++ ); bad |= Make<Terrace>( id++ );
    bad |= Make<AddDimension>( id++ ); bad |= Make<RemoveDimension>( id++ ); bad |= Make<Modulus>( id++ );
    bad |= Make<DomainRotatePlane>( id++ );
    CHECK( !bad );
    CHECK( id == 47 && (int)Metadata::GetAll().size() == 47 );

    // ---- ownership: children are held by their parents
    auto fbm = New<FractalFBm>();
    {
        auto simplex = New<Simplex>();
        fbm->SetSource( simplex );
        CHECK( simplex.use_count() == 2 );
        fbm->SetGain( New<Value>() );
    }
    SmartNode<> fbmBase = fbm;
    CHECK( !SmartNode<Simplex>::DynamicCast( fbmBase ) );
    CHECK( fbm->GetActiveFeatureSet() == FastSIMD::FeatureSet::GPU );
    // a CPU SIMD level is not served by this backend
    CHECK( !New<Perlin>( FastSIMD::FeatureSet::Scalar ) );

    // ---- encoded trees
    CHECK( !NewFromEncodedNodeTree( "not a tree" ) );
    SmartNode<> encoded = NewFromEncodedNodeTree( argv[2] );
    CHECK( encoded );
    CHECK( encoded->GetMetadata().id == 13 );   // FractalFBm

    // ---- typed setters: DomainWarpGradient (amplitude 30, frequency 0.01)
    // over CellularDistance (Euclidean, Index0Add1, jitter 1, frequency 0.02)
    auto cell = New<CellularDistance>();
    cell->SetScale( 50.0f );
    cell->SetDistanceFunction( DistanceFunction::Euclidean );
    cell->SetReturnType( CellularDistance::ReturnType::Index0Add1 );
    cell->SetDistanceIndex0( 0 );
    cell->SetDistanceIndex1( 1 );
    cell->SetGridJitter( 1.0f );
    auto warp = New<DomainWarpGradient>();
    warp->SetSource( cell );
    warp->SetScale( 100.0f );
    warp->SetWarpAmplitude( 30.0f );

    // ---- generation into host buffers
    std::vector<float> g2( 16 );
    OutputMinMax mm = fbm->GenUniformGrid2D( g2.data(), 0.0f, 0.0f, 4, 4, 1.0f, 1.0f, 1337 );
    CHECK( Save( "fbm_gain_value_grid2d_4x4", g2, mm ) );

    std::vector<float> g3( 32 * 24 * 16 );
    mm = warp->GenUniformGrid3D( g3.data(), -3.5f, 2.0f, 10.0f, 32, 24, 16, 1.0f, 1.5f, 0.75f, 1337 );
    CHECK( Save( "warp_cell_grid3d_32x24x16", g3, mm ) );

    std::vector<float> t2( 64 * 48 );
    mm = encoded->GenTileable2D( t2.data(), 64, 48, 1.0f, 1.0f, 42 );
    CHECK( Save( "fbm_tileable_64x48", t2, mm ) );

    const int n = 1000;
    std::vector<float> xs( n ), ys( n ), zs( n ), p3( n );
    for( int i = 0; i < n; i++ )
    {
        xs[i] = (float)( ( i * 37 ) % 1000 ) * 0.5f - 250.0f;
        ys[i] = (float)( ( i * 91 ) % 1000 ) * 0.25f - 125.0f;
        zs[i] = (float)( ( i * 53 ) % 1000 ) * 0.75f - 375.0f;
    }
    CHECK( Save( "positions_x", xs, OutputMinMax() ) && Save( "positions_y", ys, OutputMinMax() ) && Save( "positions_z", zs, OutputMinMax() ) );
    mm = warp->GenPositionArray3D( p3.data(), n, xs.data(), ys.data(), zs.data(), 0.5f, -1.0f, 2.0f, 1337 );
    CHECK( Save( "warp_cell_positions3d_1000", p3, mm ) );

    std::vector<float> single = { warp->GenSingle3D( 12.25f, -7.5f, 3.0f, 1337 ), fbm->GenSingle2D( 1.5f, 2.5f, 1337 ) };
    CHECK( Save( "singles", single, OutputMinMax() ) );

    std::printf( "drop-in API checks passed; outputs in %s\n", gDir.c_str() );
    return 0;
}
