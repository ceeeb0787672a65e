// Node classes of the MI355X backend: each concrete node reports its id,
// parameters and inputs to the tree compiler (Generator::Describe), plus the
// reference-count plumbing behind SmartNode.
#include "fn2_internal.h"

using namespace FastNoise;
using namespace FastNoise::Internal;

namespace FastNoise
{
    namespace Internal
    {
        int32_t NodeRefAdd( const Generator* g, int32_t add )
        {
            if( add ) return g->mReferences.fetch_add( add, std::memory_order_acq_rel );
            return g->mReferences.load( std::memory_order_acquire );
        }

        void DestroyNode( const Generator* g )
        {
            delete g;
        }

        void BumpNodeRefences( const Generator* g, bool up )
        {
            if( up ) NodeRefAdd( g, 1 );
            else if( NodeRefAdd( g, -1 ) == 1 ) DestroyNode( g );
        }

        SmartNode<> ToSmartNode( const void* p )
        {
            return SmartNode<>( static_cast<Generator*>( const_cast<void*>( p ) ) );
        }

        bool FeatureSetAllowed( FastSIMD::FeatureSet f )
        {
            uint32_t v = (uint32_t)f;
            return v == 0 || v >= (uint32_t)FastSIMD::FeatureSet::GPU;
        }
    }

    void SmartNodeManager::SetMemoryPoolSize( uint32_t ) {}

    FastSIMD::FeatureSet Generator::GetActiveFeatureSet() const { return FastSIMD::FeatureSet::GPU; }
}

// Common parameter blocks of the mixins
#define DESC_SCALABLE desc.scale = mScale; desc.frequency = mFrequency;
#define DESC_SEEDED desc.seedOffset = mSeedOffset;
#define DESC_RANGE desc.rangeMin = mRangeMin; desc.rangeScale = mRangeScale;
#define DESC_CELLULAR \
    desc.i[0] = (int)mDistanceFunction; \
    desc.hyb[0] = Hyb( mMinkowskiP ); desc.hyb[1] = Hyb( mGridJitter ); desc.hyb[2] = Hyb( mSizeJitter );
#define DESC_FRACTAL \
    desc.src[0] = mSource.base.get(); desc.hyb[0] = Hyb( mGain ); desc.hyb[1] = Hyb( mWeightedStrength ); \
    desc.i[0] = mOctaves; desc.f[0] = mLacunarity;
#define DESC_WARP \
    desc.src[0] = mSource.base.get(); desc.hyb[0] = Hyb( mWarpAmplitude ); \
    for( int a = 0; a < 4; a++ ) desc.f[a] = mAxisScale[a];

void Constant::Describe( NodeDesc& desc ) const { desc.id = kConstant; desc.f[0] = mValue; }
void White::Describe( NodeDesc& desc ) const { desc.id = kWhite; DESC_SEEDED DESC_RANGE }
void Checkerboard::Describe( NodeDesc& desc ) const { desc.id = kCheckerboard; DESC_SCALABLE DESC_RANGE }
void SineWave::Describe( NodeDesc& desc ) const { desc.id = kSineWave; DESC_SCALABLE DESC_RANGE }
void Gradient::Describe( NodeDesc& desc ) const
{
    desc.id = kGradient;
    for( int a = 0; a < 4; a++ ) { desc.f[a] = mMultiplier[a]; desc.hyb[a] = Hyb( mOffset[a] ); }
}
void DistanceToPoint::Describe( NodeDesc& desc ) const
{
    desc.id = kDistanceToPoint;
    desc.i[0] = (int)mDistanceFunction;
    for( int a = 0; a < 4; a++ ) desc.hyb[a] = Hyb( mPoint[a] );
    desc.hyb[4] = Hyb( mMinkowskiP );
}
void Simplex::Describe( NodeDesc& desc ) const { desc.id = kSimplex; DESC_SCALABLE DESC_SEEDED DESC_RANGE }
void SuperSimplex::Describe( NodeDesc& desc ) const { desc.id = kSuperSimplex; DESC_SCALABLE DESC_SEEDED DESC_RANGE }
void Perlin::Describe( NodeDesc& desc ) const { desc.id = kPerlin; DESC_SCALABLE DESC_SEEDED DESC_RANGE }
void Value::Describe( NodeDesc& desc ) const { desc.id = kValue; DESC_SCALABLE DESC_SEEDED DESC_RANGE }
void CellularValue::Describe( NodeDesc& desc ) const
{
    desc.id = kCellularValue; DESC_SCALABLE DESC_SEEDED DESC_RANGE DESC_CELLULAR
    desc.i[1] = mValueIndex;
}
void CellularDistance::Describe( NodeDesc& desc ) const
{
    desc.id = kCellularDistance; DESC_SCALABLE DESC_SEEDED DESC_RANGE DESC_CELLULAR
    desc.i[1] = mDistanceIndex0; desc.i[2] = mDistanceIndex1; desc.i[3] = (int)mReturnType;
}
void CellularLookup::Describe( NodeDesc& desc ) const
{
    desc.id = kCellularLookup; DESC_SCALABLE DESC_SEEDED DESC_CELLULAR
    desc.src[0] = mLookup.base.get();
}
void FractalFBm::Describe( NodeDesc& desc ) const { desc.id = kFractalFBm; DESC_FRACTAL }
void FractalRidged::Describe( NodeDesc& desc ) const { desc.id = kFractalRidged; DESC_FRACTAL }
void PingPong::Describe( NodeDesc& desc ) const { desc.id = kPingPong; desc.src[0] = mSource.base.get(); desc.hyb[0] = Hyb( mPingPongStrength ); }
void DomainWarpGradient::Describe( NodeDesc& desc ) const { desc.id = kDomainWarpGradient; DESC_SCALABLE DESC_SEEDED DESC_WARP }
void DomainWarpSimplex::Describe( NodeDesc& desc ) const
{
    desc.id = kDomainWarpSimplex; DESC_SCALABLE DESC_SEEDED DESC_WARP
    desc.i[0] = (int)mVectorizationScheme;
}
void DomainWarpSuperSimplex::Describe( NodeDesc& desc ) const
{
    desc.id = kDomainWarpSuperSimplex; DESC_SCALABLE DESC_SEEDED DESC_WARP
    desc.i[0] = (int)mVectorizationScheme;
}
void DomainWarpFractalProgressive::Describe( NodeDesc& desc ) const { desc.id = kDomainWarpFractalProgressive; DESC_FRACTAL }
void DomainWarpFractalIndependent::Describe( NodeDesc& desc ) const { desc.id = kDomainWarpFractalIndependent; DESC_FRACTAL }
#define DESC_OP_SRC desc.src[0] = mLHS.base.get(); desc.hyb[0] = Hyb( mRHS );
#define DESC_OP_HYB desc.hyb[0] = Hyb( mLHS ); desc.hyb[1] = Hyb( mRHS );
void Add::Describe( NodeDesc& desc ) const { desc.id = kAdd; DESC_OP_SRC }
void Subtract::Describe( NodeDesc& desc ) const { desc.id = kSubtract; DESC_OP_HYB }
void Multiply::Describe( NodeDesc& desc ) const { desc.id = kMultiply; DESC_OP_SRC }
void Divide::Describe( NodeDesc& desc ) const { desc.id = kDivide; DESC_OP_HYB }
void Modulus::Describe( NodeDesc& desc ) const { desc.id = kModulus; DESC_OP_HYB }
void Min::Describe( NodeDesc& desc ) const { desc.id = kMin; DESC_OP_SRC }
void Max::Describe( NodeDesc& desc ) const { desc.id = kMax; DESC_OP_SRC }
void MinSmooth::Describe( NodeDesc& desc ) const { desc.id = kMinSmooth; DESC_OP_SRC desc.hyb[1] = Hyb( mSmoothness ); }
void MaxSmooth::Describe( NodeDesc& desc ) const { desc.id = kMaxSmooth; DESC_OP_SRC desc.hyb[1] = Hyb( mSmoothness ); }
void PowFloat::Describe( NodeDesc& desc ) const { desc.id = kPowFloat; desc.hyb[0] = Hyb( mValue ); desc.hyb[1] = Hyb( mPow ); }
void PowInt::Describe( NodeDesc& desc ) const { desc.id = kPowInt; desc.src[0] = mValue.base.get(); desc.i[0] = mPow; }
void Fade::Describe( NodeDesc& desc ) const
{
    desc.id = kFade;
    desc.src[0] = mA.base.get(); desc.src[1] = mB.base.get();
    desc.hyb[0] = Hyb( mFade ); desc.hyb[1] = Hyb( mFadeMin ); desc.hyb[2] = Hyb( mFadeMax );
    desc.i[0] = (int)mInterpolation;
}
void DomainScale::Describe( NodeDesc& desc ) const { desc.id = kDomainScale; desc.src[0] = mSource.base.get(); desc.f[0] = mScale; }
void DomainOffset::Describe( NodeDesc& desc ) const
{
    desc.id = kDomainOffset; desc.src[0] = mSource.base.get();
    for( int a = 0; a < 4; a++ ) desc.hyb[a] = Hyb( mOffset[a] );
}
void DomainRotate::Describe( NodeDesc& desc ) const
{
    desc.id = kDomainRotate; desc.src[0] = mSource.base.get();
    const float v[15] = { mYawCos, mYawSin, mPitchCos, mPitchSin, mRollCos, mRollSin, mXa, mXb, mXc, mYa, mYb, mYc, mZa, mZb, mZc };
    for( int k = 0; k < 15; k++ ) desc.f[k] = v[k];
}
void DomainAxisScale::Describe( NodeDesc& desc ) const
{
    desc.id = kDomainAxisScale; desc.src[0] = mSource.base.get();
    for( int a = 0; a < 4; a++ ) desc.f[a] = mScale[a];
}
void SeedOffset::Describe( NodeDesc& desc ) const { desc.id = kSeedOffset; desc.src[0] = mSource.base.get(); desc.i[0] = mOffset; }
void ConvertRGBA8::Describe( NodeDesc& desc ) const { desc.id = kConvertRGBA8; desc.src[0] = mSource.base.get(); desc.f[0] = mMin; desc.f[1] = mMax; }
void GeneratorCache::Describe( NodeDesc& desc ) const { desc.id = kGeneratorCache; desc.src[0] = mSource.base.get(); }
void Remap::Describe( NodeDesc& desc ) const
{
    desc.id = kRemap; desc.src[0] = mSource.base.get();
    desc.hyb[0] = Hyb( mFromMin ); desc.hyb[1] = Hyb( mFromMax ); desc.hyb[2] = Hyb( mToMin ); desc.hyb[3] = Hyb( mToMax );
    desc.i[0] = (int)mClampOutput;
}
void Terrace::Describe( NodeDesc& desc ) const
{
    desc.id = kTerrace; desc.src[0] = mSource.base.get();
    desc.f[0] = mStepCount; desc.f[1] = mStepCountRecip; desc.f[2] = mSmoothnessRecip;
    desc.hyb[0] = Hyb( mSmoothness );
}
void AddDimension::Describe( NodeDesc& desc ) const { desc.id = kAddDimension; desc.src[0] = mSource.base.get(); desc.hyb[0] = Hyb( mNewDimensionPosition ); }
void RemoveDimension::Describe( NodeDesc& desc ) const { desc.id = kRemoveDimension; desc.src[0] = mSource.base.get(); desc.i[0] = (int)mRemoveDimension; }
void SignedSquareRoot::Describe( NodeDesc& desc ) const { desc.id = kSignedSquareRoot; desc.src[0] = mSource.base.get(); }
void Abs::Describe( NodeDesc& desc ) const { desc.id = kAbs; desc.src[0] = mSource.base.get(); }
void DomainRotatePlane::Describe( NodeDesc& desc ) const { desc.id = kDomainRotatePlane; desc.src[0] = mSource.base.get(); desc.i[0] = (int)mRotationType; }
