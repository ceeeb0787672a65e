// Test infrastructure only: node defaults, the encoded-tree decoder and the
// C interface of the CPU oracle (scalar lanes). See fn2_oracle.h.
//
// The decoder restates FastNoise2's wire format: Base64 with '@' run-length
// compression of 'A' (src/FastNoise/Base64.h:94-192) and the node stream of
// src/FastNoise/Metadata.cpp:310-506. Member tables (variables, node lookups,
// hybrids, in metadata construction order) are written out per node id from
// the reference's MetadataT constructors in include/FastNoise/Generators/*.h.
#include <memory>
#include <vector>
#include <string>
#include <cstring>
#include <cmath>
#include "fn2_oracle.h"
#include "fn2_oracle_run.h"

using namespace fn2o;

namespace fn2o_simd
{
    // Lane-16 drivers built with AVX2 / AVX-512 flags (fn2_oracle_simd.cpp)
    void avx2_run( const fn2o_data::DriverArgs& a, int mode, int threads, float* minmax );
    void avx512_run( const fn2o_data::DriverArgs& a, int mode, int threads, float* minmax );
}

namespace fn2o_data
{
    Node make_node( int id )
    {
        Node n;
        n.id = id;
        for( auto& h : n.hyb ) { h.node = nullptr; h.c = 0.0f; }
        switch( id )
        {
        case kConstant: n.f[0] = 1.0f; break;
        case kGradient: break;
        case kDistanceToPoint: n.i[0] = 1; n.hyb[4].c = 1.5f; break;        // class default EuclideanSquared (BasicGenerators.h:289)
        case kCellularValue: case kCellularDistance: case kCellularLookup:
            n.i[0] = 1; n.hyb[0].c = 1.5f; n.hyb[1].c = 1.0f; n.hyb[2].c = 0.0f; // Cellular.h:23-26
            if( id == kCellularDistance ) { n.i[1] = 0; n.i[2] = 1; n.i[3] = 0; }
            break;
        case kFractalFBm: case kFractalRidged: case kDomainWarpFractalProgressive: case kDomainWarpFractalIndependent:
            n.hyb[0].c = 0.5f; n.hyb[1].c = 0.0f; n.i[0] = 3; n.f[0] = 2.0f; break; // Fractal.h:19-24
        case kPingPong: n.hyb[0].c = 2.0f; break;
        case kDomainWarpSimplex: case kDomainWarpSuperSimplex: case kDomainWarpGradient:
            n.hyb[0].c = 50.0f; n.f[0] = n.f[1] = n.f[2] = n.f[3] = 1.0f; n.i[0] = 0; break; // DomainWarp.h:18-20
        case kMinSmooth: case kMaxSmooth: n.hyb[1].c = 0.1f; break;
        case kPowFloat: n.hyb[0].c = 2.0f; n.hyb[1].c = 2.0f; break;
        case kPowInt: n.i[0] = 2; break;
        case kDomainScale: n.f[0] = 1.0f; break;
        case kDomainRotate:
            n.f[0] = 1; n.f[1] = 0; n.f[2] = 1; n.f[3] = 0; n.f[4] = 1; n.f[5] = 0;
            n.f[6] = 1; n.f[7] = 0; n.f[8] = 0; n.f[9] = 0; n.f[10] = 1; n.f[11] = 0; n.f[12] = 0; n.f[13] = 0; n.f[14] = 1;
            break;
        case kDomainAxisScale: n.f[0] = n.f[1] = n.f[2] = n.f[3] = 1.0f; break;
        case kSeedOffset: n.i[0] = 1; break;
        case kConvertRGBA8: n.f[0] = -1.0f; n.f[1] = 1.0f; break;
        case kFade: n.hyb[0].c = 0.0f; n.hyb[1].c = -1.0f; n.hyb[2].c = 1.0f; n.i[0] = 0; break;
        case kRemap: n.hyb[0].c = -1.0f; n.hyb[1].c = 1.0f; n.hyb[2].c = 0.0f; n.hyb[3].c = 1.0f; n.i[0] = 0; break;
        case kTerrace: n.f[0] = 1.0f; n.f[1] = 1.0f; n.f[2] = 0.0f; break;
        case kRemoveDimension: n.i[0] = 1; break;
        default: break;
        }
        return n;
    }
}

namespace
{
    // ---------------- member tables ----------------
    using SetVar = void ( * )( Node&, int32_t bits );
    struct Tables
    {
        std::vector<SetVar> vars;
        std::vector<int> lookups;   // src slot
        std::vector<int> hybrids;   // hyb slot
    };

    float F32( int32_t b ) { return float_of( (uint32_t)b ); }

    void setScale( Node& n, int32_t b ) { n.scale = F32( b ); n.frequency = 1.0f / n.scale; }     // BasicGenerators.h:20-24
    void setSeed( Node& n, int32_t b ) { n.seedOffset = b; }
    void setOutMin( Node& n, int32_t b ) { float v = F32( b ); n.rangeScale += n.rangeMin - v; n.rangeMin = v; } // :96-100
    void setOutMax( Node& n, int32_t b ) { n.rangeScale = F32( b ) - n.rangeMin; }                  // :105-108
    template<int K> void setF( Node& n, int32_t b ) { n.f[K] = F32( b ); }
    template<int K> void setI( Node& n, int32_t b ) { n.i[K] = b; }
    int clamp03( int v ) { return v < 0 ? 0 : v > 3 ? 3 : v; }
    void setIdx1( Node& n, int32_t b ) { n.i[1] = clamp03( b ); }                                  // Cellular.h:65,104
    void setIdx2( Node& n, int32_t b ) { n.i[2] = clamp03( b ); }                                  // Cellular.h:105
    void calcRotation( Node& n )                                                                    // Modifiers.h:107-121
    {
        float yc = n.f[0], ys = n.f[1], pc = n.f[2], ps = n.f[3], rc = n.f[4], rs = n.f[5];
        n.f[6] = yc * pc;
        n.f[7] = yc * ps * rs - ys * rc;
        n.f[8] = yc * ps * rc + ys * rs;
        n.f[9] = ys * pc;
        n.f[10] = ys * ps * rs + yc * rc;
        n.f[11] = ys * ps * rc - yc * rs;
        n.f[12] = -ps;
        n.f[13] = pc * rs;
        n.f[14] = pc * rc;
    }
    void setYaw( Node& n, int32_t b ) { float v = F32( b ); n.f[0] = std::cos( v ); n.f[1] = std::sin( v ); calcRotation( n ); }
    void setPitch( Node& n, int32_t b ) { float v = F32( b ); n.f[2] = std::cos( v ); n.f[3] = std::sin( v ); calcRotation( n ); }
    void setRoll( Node& n, int32_t b ) { float v = F32( b ); n.f[4] = std::cos( v ); n.f[5] = std::sin( v ); calcRotation( n ); }
    void setStepCount( Node& n, int32_t b ) { n.f[0] = F32( b ); n.f[1] = 1 / n.f[0]; }            // Modifiers.h:289

    Tables tables( int id )
    {
        Tables t;
        auto scalable = [&]() { t.vars.push_back( setScale ); };
        auto seeded = [&]() { t.vars.push_back( setSeed ); };
        auto range = [&]() { t.vars.push_back( setOutMin ); t.vars.push_back( setOutMax ); };
        auto perDimF = [&]() { t.vars.push_back( setF<0> ); t.vars.push_back( setF<1> ); t.vars.push_back( setF<2> ); t.vars.push_back( setF<3> ); };
        switch( id )
        {
        case kConstant: t.vars.push_back( setF<0> ); break;
        case kWhite: seeded(); range(); break;
        case kCheckerboard: case kSineWave: scalable(); range(); break;
        case kGradient: perDimF(); t.hybrids = { 0, 1, 2, 3 }; break;
        case kDistanceToPoint: t.vars.push_back( setI<0> ); t.hybrids = { 0, 1, 2, 3, 4 }; break;
        case kSimplex: case kSuperSimplex: case kPerlin: case kValue: scalable(); seeded(); range(); break;
        case kCellularValue:
            scalable(); seeded(); range(); t.vars.push_back( setI<0> ); t.vars.push_back( setIdx1 );
            t.hybrids = { 0, 1, 2 }; break;
        case kCellularDistance:
            scalable(); seeded(); range(); t.vars.push_back( setI<0> ); t.vars.push_back( setIdx1 ); t.vars.push_back( setIdx2 ); t.vars.push_back( setI<3> );
            t.hybrids = { 0, 1, 2 }; break;
        case kCellularLookup:
            scalable(); seeded(); t.vars.push_back( setI<0> ); t.lookups = { 0 }; t.hybrids = { 0, 1, 2 }; break;
        case kFractalFBm: case kFractalRidged: case kDomainWarpFractalProgressive: case kDomainWarpFractalIndependent:
            t.lookups = { 0 }; t.hybrids = { 0, 1 }; t.vars.push_back( setI<0> ); t.vars.push_back( setF<0> ); break;
        case kPingPong: t.lookups = { 0 }; t.hybrids = { 0 }; break;
        case kDomainWarpSimplex: case kDomainWarpSuperSimplex:
            scalable(); seeded(); perDimF(); t.vars.push_back( setI<0> ); t.lookups = { 0 }; t.hybrids = { 0 }; break;
        case kDomainWarpGradient: scalable(); seeded(); perDimF(); t.lookups = { 0 }; t.hybrids = { 0 }; break;
        case kAdd: case kMultiply: case kMin: case kMax: t.lookups = { 0 }; t.hybrids = { 0 }; break;
        case kSubtract: case kDivide: case kModulus: t.hybrids = { 0, 1 }; break;
        case kAbs: case kSignedSquareRoot: case kGeneratorCache: t.lookups = { 0 }; break;
        case kMinSmooth: case kMaxSmooth: t.lookups = { 0 }; t.hybrids = { 0, 1 }; break;
        case kPowFloat: t.hybrids = { 0, 1 }; break;
        case kPowInt: t.lookups = { 0 }; t.vars.push_back( setI<0> ); break;
        case kDomainScale: t.lookups = { 0 }; t.vars.push_back( setF<0> ); break;
        case kDomainOffset: t.lookups = { 0 }; t.hybrids = { 0, 1, 2, 3 }; break;
        case kDomainRotate: t.lookups = { 0 }; t.vars = { setYaw, setPitch, setRoll }; break;
        case kDomainAxisScale: t.lookups = { 0 }; perDimF(); break;
        case kSeedOffset: t.lookups = { 0 }; t.vars.push_back( setI<0> ); break;
        case kConvertRGBA8: t.lookups = { 0 }; t.vars.push_back( setF<0> ); t.vars.push_back( setF<1> ); break;
        case kFade: t.lookups = { 0, 1 }; t.hybrids = { 0, 1, 2 }; t.vars.push_back( setI<0> ); break;
        case kRemap: t.lookups = { 0 }; t.hybrids = { 0, 1, 2, 3 }; t.vars.push_back( setI<0> ); break;
        case kTerrace: t.lookups = { 0 }; t.vars.push_back( setStepCount ); t.hybrids = { 0 }; break;
        case kAddDimension: t.lookups = { 0 }; t.hybrids = { 0 }; break;
        case kRemoveDimension: case kDomainRotatePlane: t.lookups = { 0 }; t.vars.push_back( setI<0> ); break;
        default: break;
        }
        return t;
    }

    bool isWarp( const Node* n ) { return n && ( n->id == kDomainWarpSimplex || n->id == kDomainWarpSuperSimplex || n->id == kDomainWarpGradient ); }

    void setHybridValue( Node& n, int slot, float v )
    {
        n.hyb[slot].node = nullptr;   // assigning a float replaces the hybrid (Generator.h:143-148)
        n.hyb[slot].c = v;
        if( n.id == kTerrace && v != 0.0f ) n.f[2] = 1 + 1 / v;   // Modifiers.h:291
    }

    struct Tree
    {
        std::vector<std::unique_ptr<Node>> nodes;
        const Node* root = nullptr;
    };

    // ---------------- Base64 (Base64.h:94-192) ----------------
    int b64val( unsigned char c )
    {
        if( c >= 'A' && c <= 'Z' ) return c - 'A';
        if( c >= 'a' && c <= 'z' ) return c - 'a' + 26;
        if( c >= '0' && c <= '9' ) return c - '0' + 52;
        if( c == '+' ) return 62;
        if( c == '/' ) return 63;
        if( c == '=' ) return 0;
        return 64;
    }

    bool b64decode( const char* in, std::vector<uint8_t>& out )
    {
        // expand '@<n>' into n+3 'A' characters, then decode 4 characters per 3 bytes
        std::string s;
        for( size_t i = 0; in[i]; )
        {
            if( in[i] == '@' )
            {
                int extra = b64val( (unsigned char)in[i + 1] );
                if( in[i + 1] == '\0' || extra == 64 ) return false;
                s.append( (size_t)extra + 3, 'A' );
                i += 2;
            }
            else s += in[i++];
        }
        if( s.empty() || s.size() % 4 != 0 ) return false;
        size_t outLen = s.size() / 4 * 3;
        // the reference tests the trailing '=' of the raw (compressed) input
        size_t rawLen = strlen( in );
        if( in[rawLen - 1] == '=' ) { outLen--; if( rawLen >= 2 && in[rawLen - 2] == '=' ) outLen--; }
        out.assign( outLen, 0 );
        size_t j = 0;
        for( size_t i = 0; i < s.size(); i += 4 )
        {
            uint32_t a = b64val( (unsigned char)s[i] ), b = b64val( (unsigned char)s[i + 1] ), c = b64val( (unsigned char)s[i + 2] ), d = b64val( (unsigned char)s[i + 3] );
            uint32_t triple = ( a << 18 ) + ( b << 12 ) + ( c << 6 ) + d;
            if( j < outLen ) out[j++] = ( triple >> 16 ) & 0xFF;
            if( j < outLen ) out[j++] = ( triple >> 8 ) & 0xFF;
            if( j < outLen ) out[j++] = triple & 0xFF;
        }
        return true;
    }

    // ---------------- node stream (Metadata.cpp:310-506) ----------------
    struct Stream
    {
        std::vector<uint8_t> b;
        size_t pos = 0;
        int nodeEnds = 0;   // pending NodeEnd markers from one run-length byte

        template<typename T> bool get( T& v )
        {
            if( nodeEnds > 0 || pos + sizeof( T ) > b.size() ) return false;
            std::memcpy( &v, b.data() + pos, sizeof( T ) );
            pos += sizeof( T );
            return true;
        }
        // member lookup byte: type in the low 3 bits, index in the high 5 (Metadata.cpp:88-106)
        bool member( int& type, int& index )
        {
            if( nodeEnds > 0 ) { nodeEnds--; type = 4; index = 0; return true; }
            if( pos + 1 > b.size() ) return false;
            uint8_t m = b[pos++];
            type = m & 7;
            index = m >> 3;
            if( type == 4 ) nodeEnds = index;
            return true;
        }
    };

    Node* decodeNode( Stream& s, Tree& t, std::vector<Node*>& refs )
    {
        uint8_t id;
        if( !s.get( id ) ) return nullptr;
        if( id == 0xFF )
        {
            uint16_t ref;
            if( !s.get( ref ) || ref >= refs.size() ) return nullptr;
            return refs[ref];
        }
        if( id >= kNodeCount ) return nullptr;
        t.nodes.emplace_back( new Node( make_node( id ) ) );
        Node* n = t.nodes.back().get();
        Tables tb = tables( id );
        int type, index;
        if( !s.member( type, index ) ) return nullptr;
        while( type == 0 )   // variables
        {
            int32_t v;
            if( !s.get( v ) ) return nullptr;
            if( (size_t)index < tb.vars.size() ) tb.vars[index]( *n, v );
            if( !s.member( type, index ) ) return nullptr;
        }
        if( type == 1 )      // node lookups, count in index
        {
            size_t i = 0;
            for( ; i < std::min<size_t>( index, tb.lookups.size() ); i++ )
            {
                Node* c = decodeNode( s, t, refs );
                if( !c ) return nullptr;
                if( ( id == kDomainWarpFractalProgressive || id == kDomainWarpFractalIndependent ) && !isWarp( c ) ) return nullptr;
                n->src[tb.lookups[i]] = c;
            }
            for( ; i < (size_t)index; i++ )
                if( !decodeNode( s, t, refs ) ) return nullptr;
            for( ; i < tb.lookups.size(); i++ )
            {
                // missing lookups are filled with a Constant node (Metadata.cpp:443-450)
                if( id == kDomainWarpFractalProgressive || id == kDomainWarpFractalIndependent ) return nullptr;
                t.nodes.emplace_back( new Node( make_node( kConstant ) ) );
                n->src[tb.lookups[i]] = t.nodes.back().get();
            }
            if( !s.member( type, index ) ) return nullptr;
        }
        while( type != 4 )   // hybrids until node end
        {
            if( type == 2 )
            {
                Node* c = decodeNode( s, t, refs );
                if( (size_t)index < tb.hybrids.size() )
                {
                    if( !c ) return nullptr;
                    n->hyb[tb.hybrids[index]].node = c;
                }
            }
            else if( type == 3 )
            {
                float v;
                if( !s.get( v ) ) return nullptr;
                if( (size_t)index < tb.hybrids.size() ) setHybridValue( *n, tb.hybrids[index], v );
            }
            if( !s.member( type, index ) ) return nullptr;
        }
        refs.push_back( n );
        return n;
    }

    const Tree* T( const void* p ) { return static_cast<const Tree*>( p ); }

    int run( const DriverArgs& a, int mode, int lanes, int threads, float* minmax )
    {
        if( !a.root ) return -1;
#if !defined( FN2O_COUNTING ) && !defined( FN2O_TRACING )
        if( lanes == 16 )
        {
            __builtin_cpu_init();
            if( __builtin_cpu_supports( "avx512f" ) && __builtin_cpu_supports( "avx512dq" ) && __builtin_cpu_supports( "avx512vl" ) && __builtin_cpu_supports( "avx512bw" ) )
            {
                fn2o_simd::avx512_run( a, mode, threads, minmax );
                return 0;
            }
            if( __builtin_cpu_supports( "avx2" ) && __builtin_cpu_supports( "fma" ) )
            {
                fn2o_simd::avx2_run( a, mode, threads, minmax );
                return 0;
            }
        }
#else
        (void)lanes;
        threads = 1;
#endif
        run_mode<1>( a, mode, threads, minmax );
        return 0;
    }
}

extern "C" {

void* fn2o_new_from_encoded( const char* encoded )
{
    if( !encoded ) return nullptr;
    Stream s;
    if( !b64decode( encoded, s.b ) ) return nullptr;
    auto* t = new Tree;
    std::vector<Node*> refs;
    t->root = decodeNode( s, *t, refs );
    if( !t->root ) { delete t; return nullptr; }
    return t;
}

void fn2o_delete( void* tree ) { delete static_cast<Tree*>( tree ); }

int fn2o_root_id( const void* tree ) { return tree ? T( tree )->root->id : -1; }

int fn2o_gen_grid2d( const void* tree, float* out, float xo, float yo, int nx, int ny, float sx, float sy, int seed, int mode, int lanes, int threads, float* minmax )
{
    if( !tree || nx < 0 || ny < 0 ) return -1;
    DriverArgs a; a.kind = 0; a.dims = 2; a.root = T( tree )->root; a.out = out; a.total = (int64_t)nx * ny;
    a.count[0] = nx; a.count[1] = ny; a.offset[0] = xo; a.offset[1] = yo; a.step[0] = sx; a.step[1] = sy; a.seed = seed;
    return run( a, mode, lanes, threads, minmax );
}

int fn2o_gen_grid3d( const void* tree, float* out, float xo, float yo, float zo, int nx, int ny, int nz, float sx, float sy, float sz, int seed, int mode, int lanes, int threads, float* minmax )
{
    if( !tree || nx < 0 || ny < 0 || nz < 0 ) return -1;
    DriverArgs a; a.kind = 0; a.dims = 3; a.root = T( tree )->root; a.out = out; a.total = (int64_t)nx * ny * nz;
    a.count[0] = nx; a.count[1] = ny; a.count[2] = nz; a.offset[0] = xo; a.offset[1] = yo; a.offset[2] = zo;
    a.step[0] = sx; a.step[1] = sy; a.step[2] = sz; a.seed = seed;
    return run( a, mode, lanes, threads, minmax );
}

int fn2o_gen_grid4d( const void* tree, float* out, float xo, float yo, float zo, float wo, int nx, int ny, int nz, int nw, float sx, float sy, float sz, float sw, int seed, int mode, int lanes, int threads, float* minmax )
{
    if( !tree || nx < 0 || ny < 0 || nz < 0 || nw < 0 ) return -1;
    DriverArgs a; a.kind = 0; a.dims = 4; a.root = T( tree )->root; a.out = out; a.total = (int64_t)nx * ny * nz * nw;
    a.count[0] = nx; a.count[1] = ny; a.count[2] = nz; a.count[3] = nw;
    a.offset[0] = xo; a.offset[1] = yo; a.offset[2] = zo; a.offset[3] = wo;
    a.step[0] = sx; a.step[1] = sy; a.step[2] = sz; a.step[3] = sw; a.seed = seed;
    return run( a, mode, lanes, threads, minmax );
}

int fn2o_gen_pos2d( const void* tree, float* out, int count, const float* xs, const float* ys, float xo, float yo, int seed, int mode, int lanes, int threads, float* minmax )
{
    if( !tree || count < 0 ) return -1;
    DriverArgs a; a.kind = 1; a.dims = 2; a.root = T( tree )->root; a.out = out; a.total = count;
    a.posArrays[0] = xs; a.posArrays[1] = ys; a.offset[0] = xo; a.offset[1] = yo; a.seed = seed;
    return run( a, mode, lanes, threads, minmax );
}

int fn2o_gen_pos3d( const void* tree, float* out, int count, const float* xs, const float* ys, const float* zs, float xo, float yo, float zo, int seed, int mode, int lanes, int threads, float* minmax )
{
    if( !tree || count < 0 ) return -1;
    DriverArgs a; a.kind = 1; a.dims = 3; a.root = T( tree )->root; a.out = out; a.total = count;
    a.posArrays[0] = xs; a.posArrays[1] = ys; a.posArrays[2] = zs; a.offset[0] = xo; a.offset[1] = yo; a.offset[2] = zo; a.seed = seed;
    return run( a, mode, lanes, threads, minmax );
}

int fn2o_gen_pos4d( const void* tree, float* out, int count, const float* xs, const float* ys, const float* zs, const float* ws, float xo, float yo, float zo, float wo, int seed, int mode, int lanes, int threads, float* minmax )
{
    if( !tree || count < 0 ) return -1;
    DriverArgs a; a.kind = 1; a.dims = 4; a.root = T( tree )->root; a.out = out; a.total = count;
    a.posArrays[0] = xs; a.posArrays[1] = ys; a.posArrays[2] = zs; a.posArrays[3] = ws;
    a.offset[0] = xo; a.offset[1] = yo; a.offset[2] = zo; a.offset[3] = wo; a.seed = seed;
    return run( a, mode, lanes, threads, minmax );
}

int fn2o_gen_tileable2d( const void* tree, float* out, int nx, int ny, float sx, float sy, int seed, int mode, int lanes, int threads, float* minmax )
{
    if( !tree || nx < 0 || ny < 0 ) return -1;
    DriverArgs a; a.kind = 2; a.dims = 4; a.root = T( tree )->root; a.out = out; a.total = (int64_t)nx * ny;
    a.count[0] = nx; a.count[1] = ny; a.step[0] = sx; a.step[1] = sy; a.seed = seed;
    return run( a, mode, lanes, threads, minmax );
}

static float single( const void* tree, int dims, const float* p, int seed, int mode )
{
    const Node* r = T( tree )->root;
    switch( ( mode & 3 ) * 3 + ( dims - 2 ) )
    {
    case 0: return Run<1, true, false>::single<2>( r, p, seed );
    case 1: return Run<1, true, false>::single<3>( r, p, seed );
    case 2: return Run<1, true, false>::single<4>( r, p, seed );
    case 3: return Run<1, false, false>::single<2>( r, p, seed );
    case 4: return Run<1, false, false>::single<3>( r, p, seed );
    case 5: return Run<1, false, false>::single<4>( r, p, seed );
    case 6: return Run<1, true, true>::single<2>( r, p, seed );
    case 7: return Run<1, true, true>::single<3>( r, p, seed );
    case 8: return Run<1, true, true>::single<4>( r, p, seed );
    case 9: return Run<1, false, true>::single<2>( r, p, seed );
    case 10: return Run<1, false, true>::single<3>( r, p, seed );
    default: return Run<1, false, true>::single<4>( r, p, seed );
    }
}

float fn2o_gen_single2d( const void* tree, float x, float y, int seed, int mode ) { float p[2] = { x, y }; return single( tree, 2, p, seed, mode ); }
float fn2o_gen_single3d( const void* tree, float x, float y, float z, int seed, int mode ) { float p[3] = { x, y, z }; return single( tree, 3, p, seed, mode ); }
float fn2o_gen_single4d( const void* tree, float x, float y, float z, float w, int seed, int mode ) { float p[4] = { x, y, z, w }; return single( tree, 4, p, seed, mode ); }

int fn2o_hash_primes( int seed, int n, const int* p )
{
    using E = Eval<1, true, false>;
    auto s = isplat<1>( seed );
    switch( n )
    {
    case 2: return E::hashPrimes( s, isplat<1>( p[0] ), isplat<1>( p[1] ) ).v[0];
    case 3: return E::hashPrimes( s, isplat<1>( p[0] ), isplat<1>( p[1] ), isplat<1>( p[2] ) ).v[0];
    default: return E::hashPrimes( s, isplat<1>( p[0] ), isplat<1>( p[1] ), isplat<1>( p[2] ), isplat<1>( p[3] ) ).v[0];
    }
}

int fn2o_hash_primes_hb( int seed, int n, const int* p )
{
    using E = Eval<1, true, false>;
    auto s = isplat<1>( seed );
    switch( n )
    {
    case 2: return E::hashPrimesHB( s, isplat<1>( p[0] ), isplat<1>( p[1] ) ).v[0];
    case 3: return E::hashPrimesHB( s, isplat<1>( p[0] ), isplat<1>( p[1] ), isplat<1>( p[2] ) ).v[0];
    default: return E::hashPrimesHB( s, isplat<1>( p[0] ), isplat<1>( p[1] ), isplat<1>( p[2] ), isplat<1>( p[3] ) ).v[0];
    }
}

float fn2o_value_coord( int seed, int n, const int* p )
{
    using E = Eval<1, true, false>;
    auto s = isplat<1>( seed );
    switch( n )
    {
    case 2: return E::valueCoord( s, isplat<1>( p[0] ), isplat<1>( p[1] ) ).v[0];
    case 3: return E::valueCoord( s, isplat<1>( p[0] ), isplat<1>( p[1] ), isplat<1>( p[2] ) ).v[0];
    default: return E::valueCoord( s, isplat<1>( p[0] ), isplat<1>( p[1] ), isplat<1>( p[2] ), isplat<1>( p[3] ) ).v[0];
    }
}

float fn2o_grad_dot( int dims, int hash, const float* f, int mode )
{
    auto h = isplat<1>( hash );
    auto F = []( float v ) { return fsplat<1>( v ); };
    if( dims == 2 )
        return ( mode & FN2O_AVX2 ) ? Eval<1, true, true>::gradDot2( h, F( f[0] ), F( f[1] ) ).v[0]
                                    : Eval<1, true, false>::gradDot2( h, F( f[0] ), F( f[1] ) ).v[0];
    if( dims == 3 ) return Eval<1, true, false>::gradDot3( h, F( f[0] ), F( f[1] ), F( f[2] ) ).v[0];
    return Eval<1, true, false>::gradDot4( h, F( f[0] ), F( f[1] ), F( f[2] ), F( f[3] ) ).v[0];
}

float fn2o_interp( int kind, float t, int mode )
{
    auto x = fsplat<1>( t );
    if( mode & FN2O_UNFUSED )
        return kind == 0 ? Eval<1, false, false>::interpHermite( x ).v[0] : Eval<1, false, false>::interpQuintic( x ).v[0];
    return kind == 0 ? Eval<1, true, false>::interpHermite( x ).v[0] : Eval<1, true, false>::interpQuintic( x ).v[0];
}

void fn2o_warp_vec( int dims, int scheme, int hash, const float* f, float mult, float* v, int mode )
{
    using F1 = FV<1>;
    F1 fv[4], vv[4];
    for( int a = 0; a < dims; a++ ) { fv[a] = fsplat<1>( f[a] ); vv[a] = fsplat<1>( v[a] ); }
    auto h = isplat<1>( hash );
    auto m = fsplat<1>( mult );
    auto run = [&]( auto eval ) {
        using E = decltype( eval );
        if( dims == 2 ) E::template vecContribution<2>( scheme, h, fv, m, vv );
        else if( dims == 3 ) E::template vecContribution<3>( scheme, h, fv, m, vv );
        else E::template vecContribution<4>( scheme, h, fv, m, vv );
    };
    if( mode & FN2O_UNFUSED ) run( Eval<1, false, false>() );
    else run( Eval<1, true, false>() );
    for( int a = 0; a < dims; a++ ) v[a] = vv[a].v[0];
}

int fn2o_base64_decode( const char* in, unsigned char* out, int cap )
{
    std::vector<uint8_t> v;
    if( !in || !b64decode( in, v ) ) return -1;
    for( size_t i = 0; i < v.size() && (int)i < cap; i++ ) out[i] = v[i];
    return (int)v.size();
}

#ifdef FN2O_COUNTING
// Op counts of the calls made since the last reset (this thread).
void fn2o_counters( unsigned long long* flop, unsigned long long* valu )
{
    *flop = fn2o_count::flop;
    *valu = fn2o_count::valu;
}

void fn2o_reset_counters( void )
{
    fn2o_count::flop = 0;
    fn2o_count::valu = 0;
}
#endif

#ifdef FN2O_TRACING
// Index-replay build: the next generation calls of this thread write one
// decision fingerprint per point into fingerprints (null stops it).
void fn2o_set_trace( unsigned long long* fingerprints )
{
    fn2o_trace::out = reinterpret_cast<uint64_t*>( fingerprints );
}
#endif

} // extern "C"

#ifdef FN2O_TRACING
namespace fn2o_trace
{
    thread_local uint64_t fp = 0;
    thread_local uint64_t* out = nullptr;
}
#endif

#ifdef FN2O_COUNTING
namespace fn2o_count
{
    thread_local uint64_t flop = 0, valu = 0;
}
#endif
