// Node metadata registry, API-compatible with FastNoise2's
// include/FastNoise/Metadata.h: node ids in registration order, member
// tables (variables, node lookups, hybrids) with names, defaults and enum
// names, node creation by id, and the NodeData form of the wire format.
#pragma once
#include <functional>
#include <vector>
#include <string>
#include <cstdint>
#include <memory>

#include "Utility/Config.h"

namespace FastNoise
{
    class Generator;
    struct Metadata;
    struct NodeData;

    namespace Impl
    {
        template<typename T>
        FASTNOISE_API const Metadata& GetMetadata();
    }

    struct FASTNOISE_API Metadata
    {
        struct NameDesc
        {
            const char* name;
            const char* desc;

            NameDesc( const char* name, const char* desc = "" ) : name( name ), desc( desc ) {}
        };

        struct Member
        {
            const char* name = "";
            const char* description = "";
            int dimensionIdx = -1;
        };

        struct MemberVariable : Member
        {
            enum eType
            {
                EFloat,
                EInt,
                EEnum
            };

            union ValueUnion
            {
                float f;
                int i;

                ValueUnion( float v = 0 ) { f = v; }
                ValueUnion( int v ) { i = v; }

                operator float() { return f; }
                operator int() { return i; }

                bool operator==( const ValueUnion& rhs ) const { return i == rhs.i; }
            };

            eType type = EFloat;
            ValueUnion valueDefault, valueMin, valueMax;
            float valueUiDragSpeed = 0;
            std::vector<const char*> enumNames;

            std::function<bool( Generator*, ValueUnion )> setFunc;
        };

        struct MemberNodeLookup : Member
        {
            std::function<bool( Generator*, SmartNodeArg<> )> setFunc;
        };

        struct MemberHybrid : Member
        {
            float valueDefault = 0;
            float valueUiDragSpeed = 0;

            std::function<bool( Generator*, float )> setValueFunc;
            std::function<bool( Generator*, SmartNodeArg<> )> setNodeFunc;
        };

        using node_id = uint8_t;

        virtual ~Metadata() = default;

        static const std::vector<const Metadata*>& GetAll();

        static const Metadata* GetFromId( node_id nodeId );

        template<typename T>
        static const Metadata& Get()
        {
            static_assert( std::is_base_of<Generator, T>::value, "This function should only be used for FastNoise node classes, for example FastNoise::Simplex" );
            return Impl::GetMetadata<T>();
        }

        static std::string SerialiseNodeData( NodeData* nodeData, bool fixUp = false );

        static NodeData* DeserialiseNodeData( const char* serialisedBase64NodeData, std::vector<std::unique_ptr<NodeData>>& nodeDataOut );

        static std::string FormatMetadataNodeName( const Metadata* metadata, bool removeGroups = false );

        static std::string FormatMetadataMemberName( const Member& member );

        virtual SmartNode<> CreateNode( FastSIMD::FeatureSet maxFeatureSet = FastSIMD::FeatureSet::Max ) const = 0;

        node_id id = 0;
        std::vector<MemberVariable> memberVariables;
        std::vector<MemberNodeLookup> memberNodeLookups;
        std::vector<MemberHybrid> memberHybrids;
        std::vector<const char*> groups;

        const char* name = "";
        const char* description = "";
        const char* formattedName = nullptr;
    };

    // Editable form of one node of an encoded tree (Metadata.h:302-346 of the reference)
    struct NodeData
    {
        NodeData( const Metadata* data )
        {
            if( ( metadata = data ) )
            {
                for( const Metadata::MemberVariable& value : metadata->memberVariables )
                {
                    variables.push_back( value.valueDefault );
                }
                nodeLookups.assign( metadata->memberNodeLookups.size(), nullptr );
                for( const Metadata::MemberHybrid& value : metadata->memberHybrids )
                {
                    hybrids.emplace_back( nullptr, value.valueDefault );
                }
            }
        }

        const Metadata* metadata;
        std::vector<Metadata::MemberVariable::ValueUnion> variables;
        std::vector<NodeData*> nodeLookups;
        std::vector<std::pair<NodeData*, float>> hybrids;

        bool operator==( const NodeData& rhs ) const
        {
            return metadata == rhs.metadata &&
                variables == rhs.variables &&
                nodeLookups == rhs.nodeLookups &&
                hybrids == rhs.hybrids;
        }
    };
}
