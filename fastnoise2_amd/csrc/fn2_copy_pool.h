// Multi-threaded host memcpy shared by the serving engine (retired batches
// into callers' buffers) and the synchronous host-buffer entry points (their
// pinned staging ring).
#pragma once
#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace fn2
{
    // Host-side copies between pinned staging memory and callers' pageable
    // buffers, spread over a few threads: one thread's memcpy bandwidth would
    // otherwise bound the copies at roughly 10 GB/s. Run() is for one caller at a
    // time.
    class CopyPool
    {
    public:
        struct Seg { float* dst; const float* src; size_t n; };

        explicit CopyPool( int threads )
        {
            for( int t = 0; t < threads; t++ ) workers.emplace_back( [this, t] { Work( t ); } );
        }
        ~CopyPool()
        {
            {
                std::lock_guard<std::mutex> lk( m );
                quit = true;
                gen++;
            }
            cv.notify_all();
            for( auto& w : workers ) w.join();
        }
        // copies every segment; returns when all are done
        void Run( const std::vector<Seg>& segs )
        {
            size_t total = 0;
            for( const Seg& g : segs ) total += g.n;
            if( total < ( 1u << 18 ) || workers.empty() )
            {
                for( const Seg& g : segs ) memcpy( g.dst, g.src, g.n * 4 );
                return;
            }
            // split into equal float ranges, one per worker plus the caller
            parts.assign( workers.size() + 1, {} );
            size_t per = ( total + parts.size() - 1 ) / parts.size();
            size_t p = 0, fill = 0;
            for( const Seg& g : segs )
            {
                size_t off = 0;
                while( off < g.n )
                {
                    // the last part takes whatever remains
                    size_t take = p + 1 < parts.size() ? std::min( g.n - off, per - fill ) : g.n - off;
                    parts[p].push_back( { g.dst + off, g.src + off, take } );
                    off += take;
                    fill += take;
                    if( fill == per && p + 1 < parts.size() ) { p++; fill = 0; }
                }
            }
            {
                std::lock_guard<std::mutex> lk( m );
                pending = (int)workers.size();
                gen++;
            }
            cv.notify_all();
            for( const Seg& g : parts.back() ) memcpy( g.dst, g.src, g.n * 4 );
            std::unique_lock<std::mutex> lk( m );
            cvDone.wait( lk, [&] { return pending == 0; } );
        }

    private:
        void Work( int t )
        {
            uint64_t seen = 0;
            for( ;; )
            {
                {
                    std::unique_lock<std::mutex> lk( m );
                    cv.wait( lk, [&] { return gen != seen; } );
                    seen = gen;
                    if( quit ) return;
                }
                for( const Seg& g : parts[t] ) memcpy( g.dst, g.src, g.n * 4 );
                {
                    std::lock_guard<std::mutex> lk( m );
                    pending--;
                }
                cvDone.notify_all();
            }
        }

        std::vector<std::thread> workers;
        std::vector<std::vector<Seg>> parts;
        std::mutex m;
        std::condition_variable cv, cvDone;
        uint64_t gen = 0;
        int pending = 0;
        bool quit = false;
    };
}
