// Intrusive reference-counted node handle, API-compatible with FastNoise2's
// FastNoise::SmartNode<T> (include/FastNoise/Utility/SmartNode.h of the
// reference): copy/move/convert between node types, DynamicCast, get(),
// use_count(). The count lives in the Generator base object.
#pragma once
#include <cstddef>
#include <cstdint>
#include <utility>
#include <type_traits>
#include <functional>

#include "Config.h"

namespace FastNoise
{
    class FASTNOISE_API SmartNodeManager
    {
    public:
        static constexpr uint64_t kInvalidReferenceId = (uint64_t)0;

        SmartNodeManager() = delete;

        // Kept for API compatibility: nodes are allocated individually here.
        static void SetMemoryPoolSize( uint32_t size );
    };

    namespace Internal
    {
        FASTNOISE_API SmartNode<> ToSmartNode( const void* );
        FASTNOISE_API void BumpNodeRefences( const Generator*, bool up );
        FASTNOISE_API int32_t NodeRefAdd( const Generator*, int32_t add );
        FASTNOISE_API void DestroyNode( const Generator* );
    }

    template<typename T>
    class SmartNode
    {
    public:
        template<typename U>
        static SmartNode DynamicCast( const SmartNode<U>& node )
        {
            if( T* cast = dynamic_cast<T*>( node.mPtr ) )
            {
                return SmartNode<T>( cast );
            }
            return nullptr;
        }

        constexpr SmartNode( std::nullptr_t = nullptr ) noexcept : mPtr( nullptr ) {}

        SmartNode( const SmartNode& node ) noexcept : mPtr( node.mPtr ) { TryInc( mPtr ); }

        template<typename U>
        SmartNode( const SmartNode<U>& node ) noexcept : mPtr( node.mPtr ) { TryInc( mPtr ); }

        template<typename U>
        SmartNode( const SmartNode<U>&, T* ptr ) noexcept : mPtr( ptr ) { TryInc( mPtr ); }

        SmartNode( SmartNode&& node ) noexcept : mPtr( node.mPtr ) { node.mPtr = nullptr; }

        template<typename U>
        SmartNode( SmartNode<U>&& node ) noexcept : mPtr( node.mPtr ) { node.mPtr = nullptr; }

        ~SmartNode() { Release(); }

        SmartNode& operator=( SmartNode&& node ) noexcept
        {
            swap( node );
            return *this;
        }

        template<typename U>
        SmartNode& operator=( SmartNode<U>&& node ) noexcept
        {
            if( mPtr != node.mPtr )
            {
                Release();
                mPtr = node.mPtr;
                node.mPtr = nullptr;
            }
            return *this;
        }

        SmartNode& operator=( const SmartNode& node ) noexcept
        {
            if( mPtr != node.mPtr )
            {
                TryInc( node.mPtr );
                Release();
                mPtr = node.mPtr;
            }
            return *this;
        }

        template<typename U>
        SmartNode& operator=( const SmartNode<U>& node ) noexcept
        {
            if( mPtr != node.mPtr )
            {
                TryInc( node.mPtr );
                Release();
                mPtr = node.mPtr;
            }
            return *this;
        }

        template<typename U>
        friend bool operator==( const SmartNode& lhs, const SmartNode<U>& rhs ) noexcept { return lhs.get() == rhs.get(); }

        template<typename U>
        friend bool operator!=( const SmartNode& lhs, const SmartNode<U>& rhs ) noexcept { return lhs.get() != rhs.get(); }

        const T& operator*() const noexcept { return *mPtr; }
        T& operator*() noexcept { return *mPtr; }
        const T* operator->() const noexcept { return mPtr; }
        T* operator->() noexcept { return mPtr; }

        explicit operator bool() const noexcept { return mPtr != nullptr; }

        const T* get() const noexcept { return mPtr; }
        T* get() noexcept { return mPtr; }

        void reset( T* ptr = nullptr )
        {
            *this = SmartNode( ptr );
        }

        void swap( SmartNode& node ) noexcept { std::swap( mPtr, node.mPtr ); }

        long use_count() const noexcept
        {
            return mPtr ? Internal::NodeRefAdd( mPtr, 0 ) : 0;
        }

        bool unique() const noexcept { return use_count() == 1; }

    private:
        template<typename U>
        friend SmartNode<U> New( FastSIMD::FeatureSet );

        template<typename U>
        friend class SmartNode;

        friend SmartNode<> Internal::ToSmartNode( const void* );
        friend struct Metadata;

        explicit SmartNode( T* ptr ) : mPtr( ptr ) { TryInc( ptr ); }

        void Release()
        {
            if( mPtr )
            {
                if( Internal::NodeRefAdd( mPtr, -1 ) == 1 )
                {
                    Internal::DestroyNode( mPtr );
                }
            }
            mPtr = nullptr;
        }

        template<typename U>
        static void TryInc( U* ptr ) noexcept
        {
            if( ptr ) Internal::NodeRefAdd( ptr, 1 );
        }

        T* mPtr;
    };
}

namespace std
{
    template<typename T>
    struct hash<FastNoise::SmartNode<T>>
    {
        size_t operator()( const FastNoise::SmartNode<T>& node ) const noexcept
        {
            return std::hash<const T*>()( node.get() );
        }
    };
}
