// The work-queue partition of the persistent kernels (fn2::queuePart,
// fastnoise2_amd/csrc/fn2_kernel_args.h) run on the host: every wave of a
// launch runs its first item and then pulls from its part's counter until
// the part is done, the waves advancing in a random interleaving (the
// counters hand out each value once, whatever the order). Checks that every
// item runs exactly once, that every wave leaves after at most its part's
// items + 1 pulls, and that no wave waits for another, for grids of 1 to 70
// waves (fewer waves than counters included) and several item counts.
#include <cstdio>
#include <cstdint>
#include <random>
#include <vector>
#include "fn2_kernel_args.h"

static int Simulate( uint32_t waves, uint32_t items, std::mt19937& rng )
{
    std::vector<uint32_t> counter( fn2::kQueueCounters, 0 ), runs( items, 0 );
    struct Wave { uint32_t w; bool pulling, done; uint32_t pulls; fn2::QueuePart q; };
    std::vector<Wave> ws( waves );
    for( uint32_t g = 0; g < waves; g++ ) ws[g] = { g, false, g >= items, 0, fn2::queuePart( g, waves, items ) };
    uint32_t live = 0;
    for( auto& w : ws ) live += !w.done;
    while( live )
    {
        Wave& w = ws[rng() % waves];
        if( w.done ) continue;
        if( !w.pulling )
        {
            runs[w.w]++;   // the item in hand (the loop's body)
            w.pulling = true;
            if( w.q.first >= w.q.end ) { w.done = true; live--; }
            continue;
        }
        // one pull: an atomic add on the part's counter
        w.pulls++;
        // a wave leaves after at most its part's items + 1 pulls
        const uint32_t partItems = w.q.end > w.q.first ? w.q.end - w.q.first : 0;
        if( w.pulls > partItems + 1 ) { std::printf( "waves %u items %u: a wave pulls past its part\n", waves, items ); return 1; }
        uint32_t next = w.q.first + counter[w.q.part]++;
        if( next >= w.q.end ) { w.done = true; live--; continue; }
        w.w = next;
        w.pulling = false;
    }
    for( uint32_t i = 0; i < items; i++ )
    {
        if( runs[i] != 1 )
        {
            std::printf( "waves %u items %u: item %u ran %u times\n", waves, items, i, runs[i] );
            return 1;
        }
    }
    return 0;
}

int main()
{
    std::mt19937 rng( 12345 );
    int checked = 0;
    for( uint32_t waves = 1; waves <= 70; waves++ )
    {
        for( uint32_t items = 0; items <= 400; items += ( items < 80 ? 1 : 37 ) )
        {
            for( int rep = 0; rep < 3; rep++ )
            {
                if( Simulate( waves, items, rng ) ) return 1;
                checked++;
            }
        }
    }
    const uint32_t big[][2] = { { 1024, 100000 }, { 4096, 4096 }, { 4096, 4097 }, { 2048, 65536 * 4 }, { 15, 100000 }, { 17, 99999 } };
    for( auto& b : big )
    {
        if( Simulate( b[0], b[1], rng ) ) return 1;
        checked++;
    }
    std::printf( "queue partition: %d launches simulated, every item once\n", checked );
    return 0;
}
