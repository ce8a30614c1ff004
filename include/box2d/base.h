// SPDX-FileCopyrightText: 2022 Erin Catto (Box2D; the declarations below follow its public headers)
// SPDX-License-Identifier: MIT
// box2d_amd public header: base definitions.
//
// Drop-in counterpart of the reference's include/box2d/base.h (export macros, assert/log hooks,
// version, ticks and the djb2 hash used by determinism checks). Declarations keep the reference
// names and signatures so existing C code compiles unchanged against this library.
#pragma once

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#define B2_NULL_INDEX ( -1 )

#if defined( _WIN32 )
#define BOX2D_EXPORT __declspec( dllexport )
#else
#define BOX2D_EXPORT __attribute__( ( visibility( "default" ) ) )
#endif

#ifdef __cplusplus
#define B2_API extern "C" BOX2D_EXPORT
#if defined( __HIP__ )
// Compiled by hipcc: the inline math is also available to the GPU kernels.
#define B2_INLINE inline __host__ __device__
#else
#define B2_INLINE inline
#endif
#define B2_LITERAL( T ) T
#define B2_ZERO_INIT {}
#else
#define B2_API BOX2D_EXPORT
#define B2_INLINE static inline
#define B2_LITERAL( T ) ( T )
#define B2_ZERO_INIT { 0 }
#endif

/// Allocation and assert hooks (reference: include/box2d/base.h:75-98)
typedef void* b2AllocFcn( size_t size, int alignment );
typedef void b2FreeFcn( void* mem, size_t size );
typedef int b2AssertFcn( const char* condition, const char* fileName, int lineNumber );
typedef void b2LogFcn( const char* message );

B2_API void b2SetAllocator( b2AllocFcn* allocFcn, b2FreeFcn* freeFcn );
B2_API int64_t b2GetByteCount( void );
B2_API void b2SetAssertFcn( b2AssertFcn* assertFcn );
B2_API void b2SetLogFcn( b2LogFcn* logFcn );

typedef struct b2Version
{
	int major;
	int minor;
	int revision;
} b2Version;

/// Library version. This implementation tracks Box2D 3.2.0 (reference: CMakeLists.txt:4-9).
B2_API b2Version b2GetVersion( void );
B2_API bool b2IsDoublePrecision( void );

#if !defined( NDEBUG ) || defined( B2_ENABLE_ASSERT )
B2_API int b2InternalAssert( const char* condition, const char* fileName, int lineNumber );
#define B2_ASSERT( condition ) ( (void)( ( !!( condition ) ) || ( b2InternalAssert( #condition, __FILE__, (int)( __LINE__ ) ), 0 ) ) )
#else
#define B2_ASSERT( ... ) ( (void)0 )
#endif

B2_API uint64_t b2GetTicks( void );
B2_API float b2GetMilliseconds( uint64_t ticks );
B2_API float b2GetMillisecondsAndReset( uint64_t* ticks );
B2_API void b2Yield( void );

/// djb2 hash for determinism testing (reference: src/timer.c:662-671)
#define B2_HASH_INIT 5381
B2_API uint32_t b2Hash( uint32_t hash, const uint8_t* data, int count );
