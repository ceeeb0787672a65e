// Chunked-world client of the serving engine.
//
// The node editor's mesh preview (tools/NodeEditor/MeshNoisePreview.cpp of
// the reference) streams a world around a moving camera: every chunk within
// a load range is generated as a (size + 2)^3 density block with a one-voxel
// margin (BuildBloxel3DMesh :427-440), chunks are requested nearest first
// (UpdateChunksForPosition :327-395), results of a superseded generation are
// dropped by a version tag (MultiThreadQueues.h:68-111), and each block is
// culled by its value range against the iso surface before meshing
// (:442-455).
//
// Here the blocks are requests of the serving engine written into device
// buffers, and the consumer runs on the GPU as well: a chunk whose range
// straddles the iso surface gets its exposed voxel faces counted by a kernel
// (the quads a bloxel mesher would emit, :465-520); whole-air and whole-solid
// chunks are classified from the range alone. Nothing but the per-chunk
// summary crosses PCIe.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <string>
#include <set>
#include <vector>

#include "fn2_backend.h"
#include "fn2_gpu.h"

#include "fn2_launch.h"

namespace
{
    using Clock = std::chrono::steady_clock;

    // Chunk origins within the load range of a camera, nearest first
    // (UpdateChunksForPosition: the camera is moved back by half a chunk, the
    // centre chunk is the truncated position rounded toward zero to a chunk
    // multiple, and a chunk is in range when its squared integer distance is
    // at most the truncated squared range).
    std::vector<std::array<int32_t, 3>> ChunksInRange( const float cam[3], int size, float loadRange )
    {
        std::vector<std::array<int32_t, 3>> out;
        const int range = (int)std::ceil( loadRange / size );
        int32_t p[3], centre[3];
        for( int a = 0; a < 3; a++ )
        {
            p[a] = (int32_t)( cam[a] - size * 0.5f );
            centre[a] = ( p[a] / size ) * size;
        }
        const int64_t rangeSq = (int64_t)( loadRange * loadRange );
        std::vector<std::pair<int64_t, std::array<int32_t, 3>>> cand;
        for( int x = -range; x <= range; x++ )
            for( int y = -range; y <= range; y++ )
                for( int z = -range; z <= range; z++ )
                {
                    std::array<int32_t, 3> c = { x * size + centre[0], y * size + centre[1], z * size + centre[2] };
                    int64_t d2 = 0;
                    for( int a = 0; a < 3; a++ ) d2 += (int64_t)( p[a] - c[a] ) * ( p[a] - c[a] );
                    if( d2 <= rangeSq ) cand.emplace_back( d2, c );
                }
        std::stable_sort( cand.begin(), cand.end(), []( const auto& a, const auto& b ) { return a.first < b.first; } );
        out.reserve( cand.size() );
        for( auto& c : cand ) out.push_back( c.second );
        return out;
    }

    struct Pending
    {
        std::array<int32_t, 3> pos;
        int64_t request = -1;
        int buffer = -1;
        uint32_t version = 0;
        Clock::time_point tRequest;
    };

    class World
    {
    public:
        void* engine = nullptr;
        int tree = 0, size = 128, gen = 130, seed = 1337, maxInFlight = 32;
        float iso = 0.0f, loadRange = 512.0f;
        std::string initError;

        ~World()
        {
            if( stream ) ( void )hipStreamSynchronize( stream );
            // requests still in flight write into the buffers: let them finish
            for( Pending& p : inflight ) fn2ServeWait( engine, p.request, nullptr, nullptr );
            for( float* b : buffers ) if( b ) ( void )hipFree( b );
            if( dCounts ) ( void )hipFree( dCounts );
            if( hCounts ) ( void )hipHostFree( hCounts );
            if( stream ) ( void )hipStreamDestroy( stream );
        }

        bool Init()
        {
            const size_t bytes = (size_t)gen * gen * gen * sizeof( float );
            buffers.assign( maxInFlight, nullptr );
            for( float*& b : buffers )
                if( hipMalloc( (void**)&b, bytes ) != hipSuccess ) { ( void )hipGetLastError(); initError = "hipMalloc(chunk density)"; return false; }
            for( int i = maxInFlight - 1; i >= 0; i-- ) freeBuffers.push_back( i );
            if( hipMalloc( (void**)&dCounts, sizeof( uint32_t ) * 2 * maxInFlight ) != hipSuccess ||
                hipHostMalloc( (void**)&hCounts, sizeof( uint32_t ) * 2 * maxInFlight, hipHostMallocDefault ) != hipSuccess ||
                hipStreamCreateWithFlags( &stream, hipStreamNonBlocking ) != hipSuccess )
            {
                ( void )hipGetLastError();
                initError = "world device resources";
                return false;
            }
            return true;
        }

        int Update( const float cam[3] )
        {
            int submitted = 0;
            // unload what fell out of range (UpdateChunkQueues: 1.1 x the load range)
            const float unload = loadRange * 1.1f;
            int32_t p[3];
            for( int a = 0; a < 3; a++ ) p[a] = (int32_t)( cam[a] - size * 0.5f );
            for( auto it = loaded.begin(); it != loaded.end(); )
            {
                const auto& c = *it;
                int64_t d2 = 0;
                for( int a = 0; a < 3; a++ ) d2 += (int64_t)( p[a] - c[a] ) * ( p[a] - c[a] );
                if( (double)d2 > (double)unload * unload )
                {
                    registered.erase( c );
                    it = loaded.erase( it );
                }
                else ++it;
            }
            for( const auto& c : ChunksInRange( cam, size, loadRange ) )
            {
                if( freeBuffers.empty() ) break;
                if( registered.count( c ) ) continue;
                fn2GenSpec spec{};
                spec.kind = FN2_SOURCE_GRID;
                spec.dims = 3;
                for( int a = 0; a < 3; a++ )
                {
                    spec.count[a] = gen;
                    spec.offset[a] = (float)c[a] - 1.0f;   // one-voxel margin on every side
                    spec.step[a] = 1.0f;
                }
                spec.seed = seed;
                Pending pd;
                pd.pos = c;
                pd.buffer = freeBuffers.back();
                pd.version = version;
                pd.tRequest = Clock::now();
                pd.request = fn2ServeSubmit( engine, tree, &spec, buffers[pd.buffer] );
                if( pd.request < 0 ) return -1;
                freeBuffers.pop_back();
                registered.insert( c );
                inflight.push_back( pd );
                requested++;
                submitted++;
            }
            return submitted;
        }

        int Poll( fn2ChunkResult* out, int maxOut, bool waitAll )
        {
            struct Done { Pending p; float mm[2]; };
            std::vector<Done> done;
            // completed requests, oldest first; with waitAll every one in flight
            for( auto it = inflight.begin(); it != inflight.end() && (int)done.size() < maxOut; )
            {
                float mm[2];
                int st = waitAll ? ( fn2ServeWait( engine, it->request, mm, nullptr ) ? -1 : 1 )
                                 : fn2ServePoll( engine, it->request, mm, nullptr );
                if( st == 0 ) { ++it; continue; }
                if( st < 0 ) return -1;
                done.push_back( { *it, { mm[0], mm[1] } } );
                it = inflight.erase( it );
            }
            // consumer kernels for the chunks whose range straddles the iso surface
            std::vector<int> state( done.size() );
            bool launched = false;
            for( size_t i = 0; i < done.size(); i++ )
            {
                const Done& d = done[i];
                if( d.p.version != version ) continue;
                state[i] = d.mm[0] > iso ? FN2_CHUNK_AIR : d.mm[1] < iso ? FN2_CHUNK_SOLID : FN2_CHUNK_SURFACE;
                if( state[i] != FN2_CHUNK_SURFACE ) continue;
                uint32_t* counts = dCounts + 2 * d.p.buffer;
                if( !fn2::HipOk( hipMemsetAsync( counts, 0, 2 * sizeof( uint32_t ), stream ), "chunk counters" ) ||
                    !fn2::HipOk( (hipError_t)fn2_launch_bloxel_faces( buffers[d.p.buffer], gen, iso, counts, stream ), "chunk faces" ) )
                    return -1;
                launched = true;
            }
            if( launched )
            {
                if( !fn2::HipOk( hipMemcpyAsync( hCounts, dCounts, sizeof( uint32_t ) * 2 * maxInFlight, hipMemcpyDeviceToHost, stream ), "chunk counts" ) ||
                    !fn2::HipOk( hipStreamSynchronize( stream ), "chunk consumer" ) )
                    return -1;
            }
            int n = 0;
            const auto now = Clock::now();
            for( size_t i = 0; i < done.size(); i++ )
            {
                const Done& d = done[i];
                freeBuffers.push_back( d.p.buffer );
                if( d.p.version != version )
                {
                    // superseded generation (CompleteQueue::Push with an old version)
                    dropped++;
                    continue;
                }
                fn2ChunkResult& r = out[n++];
                r.x = d.p.pos[0];
                r.y = d.p.pos[1];
                r.z = d.p.pos[2];
                r.state = state[i];
                r.min = d.mm[0];
                r.max = d.mm[1];
                r.quads = state[i] == FN2_CHUNK_SURFACE ? hCounts[2 * d.p.buffer] : 0;
                r.solidVoxels = state[i] == FN2_CHUNK_SURFACE ? hCounts[2 * d.p.buffer + 1]
                              : state[i] == FN2_CHUNK_SOLID ? (uint32_t)size * size * size : 0;
                r.latencyMs = std::chrono::duration<double, std::milli>( now - d.p.tRequest ).count();
                r.version = d.p.version;
                loaded.push_back( d.p.pos );
                completed++;
            }
            return n;
        }

        uint32_t NewVersion()
        {
            version++;
            registered.clear();
            loaded.clear();
            return version;
        }

        int64_t requested = 0, completed = 0, dropped = 0;
        std::vector<std::array<int32_t, 3>> loaded;
        std::deque<Pending> inflight;

    private:
        std::vector<float*> buffers;
        std::vector<int> freeBuffers;
        std::set<std::array<int32_t, 3>> registered;
        uint32_t* dCounts = nullptr;
        uint32_t* hCounts = nullptr;
        hipStream_t stream = nullptr;
        uint32_t version = 1;
    };
}

extern "C" {

FASTNOISE_API int fn2WorldChunksInRange( const float* camera, int chunkSize, float loadRange, int32_t* chunkOrigins, int maxChunks )
{
    fn2::ClearError();
    if( !camera || chunkSize < 1 || !( loadRange >= 0.0f ) || loadRange > 65536.0f || maxChunks < 0 ||
        !( std::fabs( camera[0] ) < 1e9f && std::fabs( camera[1] ) < 1e9f && std::fabs( camera[2] ) < 1e9f ) )
    {
        fn2::SetError( "bad chunk range arguments" );
        return -1;
    }
    auto all = ChunksInRange( camera, chunkSize, loadRange );
    for( int i = 0; i < std::min<int>( maxChunks, (int)all.size() ); i++ )
        for( int a = 0; a < 3; a++ ) chunkOrigins[3 * i + a] = all[i][a];
    return (int)all.size();
}

FASTNOISE_API void* fn2WorldCreate( void* engine, int treeId, int chunkSize, float isoSurface, float loadRange, int seed, int maxInFlight )
{
    fn2::ClearError();
    // chunks up to 894: the face count of a (size + 2)^3 block stays in uint32
    if( !engine || chunkSize < 1 || chunkSize > 894 || maxInFlight < 1 || maxInFlight > 4096 ||
        !( loadRange >= 0.0f ) || loadRange > 65536.0f )
    {
        fn2::SetError( "bad world arguments" );
        return nullptr;
    }
    World* w = new World;
    w->engine = engine;
    w->tree = treeId;
    w->size = chunkSize;
    w->gen = chunkSize + 2;
    w->iso = isoSurface;
    w->loadRange = loadRange;
    w->seed = seed;
    w->maxInFlight = maxInFlight;
    if( !w->Init() )
    {
        fn2::SetError( w->initError );
        delete w;
        return nullptr;
    }
    return w;
}

FASTNOISE_API void fn2WorldDestroy( void* world )
{
    delete static_cast<World*>( world );
}

FASTNOISE_API int fn2WorldUpdate( void* world, float cameraX, float cameraY, float cameraZ )
{
    fn2::ClearError();
    if( !world ) { fn2::SetError( "null world" ); return -1; }
    if( !( std::fabs( cameraX ) < 1e9f && std::fabs( cameraY ) < 1e9f && std::fabs( cameraZ ) < 1e9f ) )
    {
        fn2::SetError( "camera position out of range" );
        return -1;
    }
    const float cam[3] = { cameraX, cameraY, cameraZ };
    return static_cast<World*>( world )->Update( cam );
}

FASTNOISE_API int fn2WorldPoll( void* world, fn2ChunkResult* results, int maxResults, int waitAll )
{
    fn2::ClearError();
    if( !world || ( !results && maxResults > 0 ) || maxResults < 0 ) { fn2::SetError( "bad poll arguments" ); return -1; }
    return static_cast<World*>( world )->Poll( results, maxResults, waitAll != 0 );
}

FASTNOISE_API uint32_t fn2WorldNewVersion( void* world )
{
    if( !world ) return 0;
    return static_cast<World*>( world )->NewVersion();
}

FASTNOISE_API int fn2WorldStats( void* world, int64_t* requested, int64_t* completed, int64_t* dropped, int* loaded, int* inFlight )
{
    if( !world ) return 1;
    World* w = static_cast<World*>( world );
    if( requested ) *requested = w->requested;
    if( completed ) *completed = w->completed;
    if( dropped ) *dropped = w->dropped;
    if( loaded ) *loaded = (int)w->loaded.size();
    if( inFlight ) *inFlight = (int)w->inflight.size();
    return 0;
}

}
