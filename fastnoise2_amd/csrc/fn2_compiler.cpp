// Tree compiler: walks a node DAG (through Generator::Describe) and emits the
// straight-line device program the interpreter kernels run.
//
// Evaluation semantics follow the reference's recursive GetSourceValue calls
// (Generator.inl:48-67): every input is evaluated where the reference
// evaluates it, with the seed and position it would receive. Seeds are
// folded: each instruction carries (node seed - call seed) as a constant.
// Children are re-emitted for every use, which unrolls fractal octaves and
// warp octaves into straight-line code.
#include <cstring>
#include <stdexcept>
#include "fn2_compiler.h"
#include "fn2_internal.h"

using namespace FastNoise;
using namespace FastNoise::Internal;

namespace fn2
{
    namespace
    {
        struct CompileError : std::runtime_error
        {
            using std::runtime_error::runtime_error;
        };

        // ScaleOutput bounds per lattice node and dimension (Perlin.inl:32,71,122;
        // Simplex.inl:60,139,283,374,574,933; Value.inl: +-2^31)
        float LatticeBound( int id, int dim )
        {
            switch( id )
            {
            case kPerlin: return dim == 2 ? 1.726796627044677734375f : dim == 3 ? (float)1.0363423824310302734375 : 1.33858621120452880859375f;
            case kSimplex: return dim == 2 ? 0.0261208079755306243896484375f : dim == 3 ? 0.030586399137973785400390625f : 0.036730043590068817138671875f;
            case kSuperSimplex: return dim == 2 ? 0.14084912836551666259765625f : dim == 3 ? 0.0069091119803488254547119140625f : (float)0.011310013942420482635498046875;
            default: return 2147483648.f;
            }
        }

        // (1 / (max - min)) * rangeScale as the reference computes it at run time
        float OutScale( float nMin, float nMax, float rangeScale )
        {
            return ( 1.0f / ( nMax - nMin ) ) * rangeScale;
        }

        uint8_t LeafOp( int id )
        {
            switch( id )
            {
            case kPerlin: return OP_PERLIN;
            case kValue: return OP_VALUE;
            case kSimplex: return OP_SIMPLEX;
            default: return OP_SUPERSIMPLEX;
            }
        }

        bool IsLattice( int id ) { return id == kPerlin || id == kValue || id == kSimplex || id == kSuperSimplex; }

        NodeDesc Desc( const Generator* g )
        {
            if( !g ) throw CompileError( "node input is not connected" );
            NodeDesc d;
            g->Describe( d );
            return d;
        }

        class Emitter
        {
        public:
            Program& P;
            uint32_t used = 0;   // register bitmap

            explicit Emitter( Program& p ) : P( p ) {}

            int Alloc( int n )
            {
                for( int b = 0; b + n <= kMaxRegs; b++ )
                {
                    uint32_t m = ( n == 32 ? 0xFFFFFFFFu : ( ( 1u << n ) - 1 ) ) << b;
                    if( ( used & m ) == 0 )
                    {
                        used |= m;
                        if( b + n > P.regsUsed ) P.regsUsed = b + n;
                        return b;
                    }
                }
                throw CompileError( "node tree needs more than 32 live registers" );
            }
            void Free( int base, int n )
            {
                if( base < 0 || base == kNoReg ) return;
                used &= ~( ( ( n == 32 ? 0xFFFFFFFFu : ( ( 1u << n ) - 1 ) ) ) << base );
            }

            uint32_t K( std::initializer_list<float> v )
            {
                uint32_t k = (uint32_t)P.konst.size();
                P.konst.insert( P.konst.end(), v.begin(), v.end() );
                return k;
            }

            Ins& Emit( uint8_t op, int dim, std::initializer_list<int> regs, int32_t seed = 0, uint32_t k = 0 )
            {
                if( P.ins.size() >= (size_t)kMaxInstructions ) throw CompileError( "node tree expands to more than 4096 instructions" );
                Ins in;
                std::memset( &in, 0, sizeof( in ) );
                in.op = op;
                in.dim = (uint8_t)dim;
                for( int i = 0; i < 6; i++ ) in.r[i] = kNoReg;
                int i = 0;
                for( int r : regs ) in.r[i++] = (uint8_t)r;
                in.seed = seed;
                in.k = k;
                P.ins.push_back( in );
                if( dim >= 2 ) P.dimMask |= 1 << dim;
                return P.ins.back();
            }

            static int32_t SeedAdd( int32_t a, int32_t b ) { return (int32_t)( (uint32_t)a + (uint32_t)b ); }

            int Const( float v )
            {
                int r = Alloc( 1 );
                Emit( OP_CONST, 1, { r }, 0, K( { v } ) );
                return r;
            }

            // Hybrid input: a fresh register holding either the node's value or the constant.
            int Hybrid( const HybridDesc& h, int32_t seed, int pos, int dim )
            {
                if( h.node ) return Gen( h.node, seed, pos, dim );
                return Const( h.c );
            }

            int CopyVec( int pos, int dim )
            {
                int r = Alloc( dim );
                Emit( OP_MOV_VEC, dim, { r, pos } );
                return r;
            }

            // Emits `g` evaluated with seed (call seed + seed) at positions
            // pos..pos+dim-1 (left unchanged). Returns a register owned by the caller.
            int Gen( const Generator* g, int32_t seed, int pos, int dim )
            {
                NodeDesc n = Desc( g );
                switch( n.id )
                {
                case kConstant: return Const( n.f[0] );

                case kPerlin: case kValue: case kSimplex: case kSuperSimplex:
                {
                    float b = LatticeBound( n.id, dim );
                    float nMin = n.id == kValue ? -2147483648.f : -b;
                    float nMax = n.id == kValue ? 2147483648.f : b;
                    int r = Alloc( 1 );
                    Emit( LeafOp( n.id ), dim, { r, pos }, SeedAdd( seed, n.seedOffset ), K( { n.frequency, nMin, OutScale( nMin, nMax, n.rangeScale ), n.rangeMin } ) );
                    return r;
                }
                case kWhite:
                {
                    P.featMask |= kFeatMisc;
                    int r = Alloc( 1 );
                    Emit( OP_WHITE, dim, { r, pos }, SeedAdd( seed, n.seedOffset ), K( { 0.0f, -2147483648.f, OutScale( -2147483648.f, 2147483648.f, n.rangeScale ), n.rangeMin } ) );
                    return r;
                }
                case kCheckerboard:
                {
                    P.featMask |= kFeatMisc;
                    int r = Alloc( 1 );
                    Emit( OP_CHECKERBOARD, dim, { r, pos }, 0, K( { n.frequency, 0.0f, OutScale( 0.0f, 2.0f, n.rangeScale ), n.rangeMin } ) );
                    return r;
                }
                case kSineWave:
                {
                    P.featMask |= kFeatMisc;
                    int r = Alloc( 1 );
                    Emit( OP_SINEWAVE, dim, { r, pos }, 0, K( { n.frequency, -1.0f, OutScale( -1.0f, 1.0f, n.rangeScale ), n.rangeMin } ) );
                    return r;
                }
                case kGradient:
                {
                    int r = Const( 0.0f );
                    for( int a = 0; a < dim; a++ )
                    {
                        int off = Hybrid( n.hyb[a], seed, pos, dim );
                        Emit( OP_GRADIENT_AXIS, 1, { r, pos + a, off }, 0, K( { n.f[a] } ) );
                        Free( off, 1 );
                    }
                    return r;
                }
                case kDistanceToPoint:
                {
                    P.featMask |= kFeatMisc;
                    int d = Alloc( dim );
                    for( int a = 0; a < dim; a++ )
                    {
                        int pt = Hybrid( n.hyb[a], seed, pos, dim );
                        Emit( OP_SUB, 1, { d + a, pos + a, pt } );
                        Free( pt, 1 );
                    }
                    int mink = n.i[0] == 5 ? Hybrid( n.hyb[4], seed, d, dim ) : kNoReg;
                    int r = Alloc( 1 );
                    Ins& in = Emit( OP_DISTANCE, dim, { r, d, mink } );
                    in.p[0] = (uint8_t)n.i[0];
                    in.p[1] = 1;
                    Free( mink, 1 );
                    Free( d, dim );
                    return r;
                }
                case kCellularValue: case kCellularDistance: case kCellularLookup:
                    return Cellular( n, seed, pos, dim );
                case kFractalFBm: case kFractalRidged:
                    return Fractal( n, seed, pos, dim );
                case kPingPong:
                {
                    int s = Gen( n.src[0], seed, pos, dim );
                    int st = Hybrid( n.hyb[0], seed, pos, dim );
                    Emit( OP_PINGPONG, 1, { s, s, st } );
                    Free( st, 1 );
                    return s;
                }
                case kDomainWarpSimplex: case kDomainWarpSuperSimplex: case kDomainWarpGradient:
                    return Warp( n, seed, pos, dim );
                case kDomainWarpFractalProgressive: case kDomainWarpFractalIndependent:
                    return WarpFractal( n, seed, pos, dim );

                case kAdd: case kMultiply: case kMin: case kMax:
                {
                    int l = Gen( n.src[0], seed, pos, dim );
                    int r = Hybrid( n.hyb[0], seed, pos, dim );
                    uint8_t op = n.id == kAdd ? OP_ADD : n.id == kMultiply ? OP_MUL : n.id == kMin ? OP_MIN : OP_MAX;
                    Emit( op, 1, { l, l, r } );
                    Free( r, 1 );
                    return l;
                }
                case kSubtract: case kDivide: case kModulus:
                {
                    int l = Hybrid( n.hyb[0], seed, pos, dim );
                    int r = Hybrid( n.hyb[1], seed, pos, dim );
                    uint8_t op = n.id == kSubtract ? OP_SUB : n.id == kDivide ? OP_DIV : OP_FMOD;
                    Emit( op, 1, { l, l, r } );
                    Free( r, 1 );
                    return l;
                }
                case kMinSmooth: case kMaxSmooth:
                {
                    int l = Gen( n.src[0], seed, pos, dim );
                    int r = Hybrid( n.hyb[0], seed, pos, dim );
                    int s = Hybrid( n.hyb[1], seed, pos, dim );
                    Ins& in = Emit( OP_MINSMOOTH, 1, { l, l, r, s } );
                    in.p[0] = n.id == kMaxSmooth;
                    Free( r, 1 );
                    Free( s, 1 );
                    return l;
                }
                case kPowFloat:
                {
                    int v = Hybrid( n.hyb[0], seed, pos, dim );
                    int p = Hybrid( n.hyb[1], seed, pos, dim );
                    Emit( OP_POWF, 1, { v, v, p } );
                    Free( p, 1 );
                    return v;
                }
                case kPowInt:
                {
                    int v = Gen( n.src[0], seed, pos, dim );
                    Ins& in = Emit( OP_POWI, 1, { v, v } );
                    in.p[0] = (uint8_t)std::min( std::max( n.i[0], 0 ), 255 );
                    return v;
                }
                case kFade:
                {
                    int a = Gen( n.src[0], seed, pos, dim );
                    int b = Gen( n.src[1], seed, pos, dim );
                    int f = Hybrid( n.hyb[0], seed, pos, dim );
                    int fmin = Hybrid( n.hyb[1], seed, pos, dim );
                    int fmax = Hybrid( n.hyb[2], seed, pos, dim );
                    Ins& in = Emit( OP_FADE, 1, { a, a, b, f, fmin, fmax } );
                    in.p[0] = (uint8_t)n.i[0];
                    Free( b, 1 ); Free( f, 1 ); Free( fmin, 1 ); Free( fmax, 1 );
                    return a;
                }
                case kAbs:
                {
                    int v = Gen( n.src[0], seed, pos, dim );
                    Emit( OP_ABS, 1, { v, v } );
                    return v;
                }
                case kSignedSquareRoot:
                {
                    int v = Gen( n.src[0], seed, pos, dim );
                    Emit( OP_SSQRT, 1, { v, v } );
                    return v;
                }
                case kGeneratorCache: return Gen( n.src[0], seed, pos, dim );
                case kSeedOffset: return Gen( n.src[0], SeedAdd( seed, n.i[0] ), pos, dim );
                case kRemap:
                {
                    int s = Gen( n.src[0], seed, pos, dim );
                    int h[4];
                    for( int a = 0; a < 4; a++ ) h[a] = Hybrid( n.hyb[a], seed, pos, dim );
                    Ins& in = Emit( OP_REMAP, 1, { s, s, h[0], h[1], h[2], h[3] } );
                    in.p[0] = (uint8_t)n.i[0];
                    for( int a = 0; a < 4; a++ ) Free( h[a], 1 );
                    return s;
                }
                case kTerrace:
                {
                    int s = Gen( n.src[0], seed, pos, dim );
                    int sm = n.hyb[0].node ? Gen( n.hyb[0].node, seed, pos, dim ) : kNoReg;
                    Ins& in = Emit( OP_TERRACE, 1, { s, s, sm }, 0, K( { n.f[0], n.f[1], n.f[2] } ) );
                    in.p[1] = n.hyb[0].c != 0.0f;
                    Free( sm, 1 );
                    return s;
                }
                case kConvertRGBA8:
                {
                    int s = Gen( n.src[0], seed, pos, dim );
                    Emit( OP_RGBA8, 1, { s, s }, 0, K( { n.f[0], n.f[1], 255.0f / ( n.f[1] - n.f[0] ) } ) );
                    return s;
                }
                case kDomainScale:
                {
                    int p = Alloc( dim );
                    Emit( OP_SCALE, dim, { p, pos }, 0, K( { n.f[0] } ) );
                    int v = Gen( n.src[0], seed, p, dim );
                    Free( p, dim );
                    return v;
                }
                case kDomainAxisScale:
                {
                    int p = Alloc( dim );
                    Emit( OP_SCALE_AXES, dim, { p, pos }, 0, K( { n.f[0], n.f[1], n.f[2], n.f[3] } ) );
                    int v = Gen( n.src[0], seed, p, dim );
                    Free( p, dim );
                    return v;
                }
                case kDomainOffset:
                {
                    int p = Alloc( dim );
                    for( int a = 0; a < dim; a++ )
                    {
                        int off = Hybrid( n.hyb[a], seed, pos, dim );
                        Emit( OP_ADD, 1, { p + a, pos + a, off } );
                        Free( off, 1 );
                    }
                    int v = Gen( n.src[0], seed, p, dim );
                    Free( p, dim );
                    return v;
                }
                case kDomainRotate:
                {
                    if( dim == 4 ) return Gen( n.src[0], seed, pos, 4 );
                    if( dim == 2 && n.f[3] == 0.0f && n.f[5] == 0.0f )
                    {
                        int p = Alloc( 2 );
                        Emit( OP_ROTATE2, 2, { p, pos }, 0, K( { n.f[0], n.f[1] } ) );
                        int v = Gen( n.src[0], seed, p, 2 );
                        Free( p, 2 );
                        return v;
                    }
                    int src = pos;
                    int tmp = kNoReg;
                    if( dim == 2 )
                    {
                        tmp = Alloc( 3 );
                        Emit( OP_MOV_VEC, 2, { tmp, pos } );
                        Emit( OP_CONST, 1, { tmp + 2 }, 0, K( { 0.0f } ) );
                        src = tmp;
                    }
                    int p = Alloc( 3 );
                    Emit( OP_ROTATE3, 3, { p, src }, 0, K( { n.f[6], n.f[7], n.f[8], n.f[9], n.f[10], n.f[11], n.f[12], n.f[13], n.f[14] } ) );
                    if( tmp != kNoReg ) Free( tmp, 3 );
                    int v = Gen( n.src[0], seed, p, 3 );
                    Free( p, 3 );
                    return v;
                }
                case kDomainRotatePlane:
                {
                    int nd = dim == 2 ? 3 : dim;
                    int p = Alloc( nd );
                    Emit( OP_MOV_VEC, dim, { p, pos } );
                    if( dim == 2 ) Emit( OP_CONST, 1, { p + 2 }, 0, K( { 0.0f } ) );
                    Ins& in = Emit( OP_ROTATE_PLANE, 3, { p, p } );
                    in.p[0] = dim == 2 ? 0 : (uint8_t)n.i[0];
                    int v = Gen( n.src[0], seed, p, nd );
                    Free( p, nd );
                    return v;
                }
                case kAddDimension:
                {
                    if( dim == 4 ) return Gen( n.src[0], seed, pos, 4 );
                    int p = Alloc( dim + 1 );
                    Emit( OP_MOV_VEC, dim, { p, pos } );
                    int nv = Hybrid( n.hyb[0], seed, pos, dim );
                    Emit( OP_MOV, 1, { p + dim, nv } );
                    Free( nv, 1 );
                    int v = Gen( n.src[0], seed, p, dim + 1 );
                    Free( p, dim + 1 );
                    return v;
                }
                case kRemoveDimension:
                {
                    int rm = n.i[0];
                    if( dim == 2 || rm < 0 || rm >= dim ) return Gen( n.src[0], seed, pos, dim );
                    int p = Alloc( dim - 1 );
                    for( int a = 0, b = 0; a < dim; a++ )
                        if( a != rm ) Emit( OP_MOV, 1, { p + b++, pos + a } );
                    int v = Gen( n.src[0], seed, p, dim - 1 );
                    Free( p, dim - 1 );
                    return v;
                }
                default:
                    throw CompileError( "unknown node type" );
                }
            }

            int Cellular( const NodeDesc& n, int32_t seed, int pos, int dim )
            {
                P.featMask |= kFeatCell;
                if( n.i[0] == 5 && n.hyb[0].node )
                    throw CompileError( "a node-valued Minkowski P on a cellular node is not supported on the GPU backend" );
                // hybrids are evaluated at the unscaled position with the un-offset seed (Cellular.inl:22-30)
                int jit = Hybrid( n.hyb[1], seed, pos, dim );
                bool sizeOn = n.hyb[2].node || n.hyb[2].c != 0.0f;
                int sj = sizeOn ? Hybrid( n.hyb[2], seed, pos, dim ) : kNoReg;
                const float jc = dim == 2 ? 0.437016f : dim == 3 ? 0.396144f : 0.366025f;
                int32_t cseed = SeedAdd( seed, n.seedOffset );
                float nMin = 0.0f, nMax = 1.0f;
                if( n.id == kCellularValue ) { nMin = -2147483648.f; nMax = 2147483648.f; }
                else if( n.id == kCellularDistance )
                {
                    float maxDist = 1 + jc;    // GetReturn (Cellular.inl:531-565)
                    maxDist *= maxDist;
                    switch( n.i[3] )
                    {
                    case 1: nMax = maxDist * 2; break;
                    case 3: nMax = maxDist * maxDist; break;
                    default: nMax = maxDist; break;
                    }
                }
                const float outScale = n.id == kCellularLookup ? 0.0f : OutScale( nMin, nMax, n.rangeScale );
                uint32_t k = K( { n.frequency, nMin, outScale, n.hyb[0].c, n.rangeMin, n.scale } );
                uint8_t op = n.id == kCellularValue ? OP_CELL_VALUE : n.id == kCellularDistance ? OP_CELL_DISTANCE : OP_CELL_LOOKUP;
                int r = n.id == kCellularLookup ? Alloc( dim ) : Alloc( 1 );
                Ins& in = Emit( op, dim, { r, pos, jit, sj }, cseed, k );
                in.p[0] = (uint8_t)n.i[0];
                in.p[1] = (uint8_t)n.i[1];
                in.p[2] = (uint8_t)n.i[2];
                in.p[3] = (uint8_t)n.i[3];
                Free( jit, 1 );
                Free( sj, 1 );
                if( n.id != kCellularLookup ) return r;
                // lookup at the cell position with seed + 1 (Cellular.inl:633, 712, 805)
                int v = Gen( n.src[0], SeedAdd( seed, 1 ), r, dim );
                Free( r, dim );
                return v;
            }

            int Fractal( const NodeDesc& n, int32_t seed, int pos, int dim )
            {
                const bool ridged = n.id == kFractalRidged;
                const int octaves = n.i[0];
                NodeDesc src = Desc( n.src[0] );
                if( IsLattice( src.id ) && !n.hyb[0].node && !n.hyb[1].node && octaves >= 1 && octaves <= 255 )
                {
                    // fused octave loop over a lattice leaf (Fractal.inl:9-67)
                    float b = LatticeBound( src.id, dim );
                    float nMin = src.id == kValue ? -2147483648.f : -b;
                    float nMax = src.id == kValue ? 2147483648.f : b;
                    int r = Alloc( 1 );
                    Ins& in = Emit( OP_FRACTAL_LEAF, dim, { r, pos }, SeedAdd( seed, src.seedOffset ),
                        K( { src.frequency, nMin, OutScale( nMin, nMax, src.rangeScale ), src.rangeMin, n.hyb[0].c, n.hyb[1].c, n.f[0] } ) );
                    in.p[0] = LeafOp( src.id );
                    in.p[1] = (uint8_t)octaves;
                    in.p[2] = ridged;
                    return r;
                }
                int gain = Hybrid( n.hyb[0], seed, pos, dim );
                int ws = Hybrid( n.hyb[1], seed, pos, dim );
                int p = CopyVec( pos, dim );
                int noise = Gen( n.src[0], seed, p, dim );
                int amp = Const( 1.0f );
                int sum = Alloc( 1 );
                if( ridged )
                {
                    Emit( OP_ABS, 1, { noise, noise } );
                    Ins& in = Emit( OP_RIDGED_TERM, 1, { sum, noise, amp } );
                    in.p[0] = 0;
                }
                else Emit( OP_MUL, 1, { sum, noise, amp } );
                for( int i = 1; i < octaves; i++ )
                {
                    Emit( ridged ? OP_RIDGED_AMP : OP_FBM_AMP, 1, { amp, noise, ws, gain } );
                    Emit( OP_SCALE, dim, { p, p }, 0, K( { n.f[0] } ) );
                    Free( noise, 1 );
                    noise = Gen( n.src[0], SeedAdd( seed, i ), p, dim );
                    if( ridged )
                    {
                        Emit( OP_ABS, 1, { noise, noise } );
                        Ins& in = Emit( OP_RIDGED_TERM, 1, { sum, noise, amp } );
                        in.p[0] = 1;
                    }
                    else Emit( OP_FBM_SUM, 1, { sum, noise, amp } );
                }
                Free( noise, 1 ); Free( amp, 1 ); Free( gain, 1 ); Free( ws, 1 ); Free( p, dim );
                return sum;
            }

            static uint8_t WarpOp( int id ) { return id == kDomainWarpGradient ? OP_WARP_GRADIENT : id == kDomainWarpSimplex ? OP_WARP_SIMPLEX : OP_WARP_SUPERSIMPLEX; }
            // seed offset added inside Warp(): gradient and simplex yes, super simplex no
            static int32_t WarpInnerSeed( const NodeDesc& w ) { return w.id == kDomainWarpSuperSimplex ? 0 : w.seedOffset; }

            // DomainWarp::GenT (DomainWarp.inl:9-17)
            int Warp( const NodeDesc& n, int32_t seed, int pos, int dim )
            {
                P.featMask |= kFeatWarp;
                int amp = Hybrid( n.hyb[0], seed, pos, dim );
                int sp = Alloc( dim );
                Emit( OP_SCALE, dim, { sp, pos }, 0, K( { n.frequency } ) );
                int out = CopyVec( pos, dim );
                int32_t wseed = SeedAdd( SeedAdd( seed, n.seedOffset ), WarpInnerSeed( n ) );
                Ins& in = Emit( WarpOp( n.id ), dim, { kNoReg, sp, out, amp }, wseed, K( { n.f[0], n.f[1], n.f[2], n.f[3] } ) );
                in.p[0] = (uint8_t)n.i[0];
                Free( sp, dim );
                Free( amp, 1 );
                int v = Gen( n.src[0], seed, out, dim );
                Free( out, dim );
                return v;
            }

            // DomainWarpFractalProgressive / Independent (DomainWarpFractal.inl:4-74)
            int WarpFractal( const NodeDesc& n, int32_t seed, int pos, int dim )
            {
                P.featMask |= kFeatWarp;
                const bool independent = n.id == kDomainWarpFractalIndependent;
                NodeDesc w = Desc( n.src[0] );
                if( w.id != kDomainWarpGradient && w.id != kDomainWarpSimplex && w.id != kDomainWarpSuperSimplex )
                    throw CompileError( "domain warp fractal source is not a domain warp" );
                int amp = Hybrid( w.hyb[0], seed, pos, dim );
                int ws = Hybrid( n.hyb[1], seed, pos, dim );
                int gain = Hybrid( n.hyb[0], seed, pos, dim );
                int freq = Const( w.frequency );
                int p = CopyVec( pos, dim );
                int sp = Alloc( dim );
                int strength = Alloc( 1 );
                const uint32_t axisK = K( { w.f[0], w.f[1], w.f[2], w.f[3] } );
                const uint32_t lacK = K( { n.f[0] } );
                for( int i = 0; i < std::max( n.i[0], 1 ); i++ )
                {
                    if( i > 0 )
                    {
                        Emit( OP_SCALE, 1, { freq, freq }, 0, lacK );
                        Emit( OP_WARP_AMP, 1, { amp, strength, ws, gain } );
                    }
                    Emit( OP_MUL_VEC, dim, { sp, independent ? pos : p, freq } );
                    Ins& in = Emit( WarpOp( w.id ), dim, { strength, sp, p, amp }, SeedAdd( SeedAdd( seed, i ), WarpInnerSeed( w ) ), axisK );
                    in.p[0] = (uint8_t)w.i[0];
                }
                Free( strength, 1 ); Free( sp, dim ); Free( freq, 1 ); Free( gain, 1 ); Free( ws, 1 ); Free( amp, 1 );
                int v = Gen( w.src[0], seed, p, dim );
                Free( p, dim );
                return v;
            }
        };

        uint64_t Fnv( const void* data, size_t n, uint64_t h )
        {
            const uint8_t* b = (const uint8_t*)data;
            for( size_t i = 0; i < n; i++ ) { h ^= b[i]; h *= 1099511628211ull; }
            return h;
        }
    }

    bool CompileTree( const Generator* root, int dims, Program& out, std::string& error )
    {
        out = Program();
        try
        {
            Emitter e( out );
            int pos = e.Alloc( dims );
            if( dims >= 2 ) out.dimMask |= 1 << dims;
            out.resultReg = e.Gen( root, 0, pos, dims );
            (void)pos;
        }
        catch( const std::exception& ex )
        {
            error = ex.what();
            return false;
        }
        uint64_t h = 1469598103934665603ull;
        h = Fnv( out.ins.data(), out.ins.size() * sizeof( Ins ), h );
        h = Fnv( out.konst.data(), out.konst.size() * sizeof( float ), h );
        h = Fnv( &out.resultReg, sizeof( int ), h );
        out.dims = dims;
        h = Fnv( &out.dims, sizeof( int ), h );
        out.hash = h;
        return true;
    }
}
