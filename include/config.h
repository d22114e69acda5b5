// config.h -- drop-in for the reference's include/config.h (compile-time
// pipeline selection).  Same macros and types; the multiplier behind them is
// libbirdwing_hip.so (gfx950).
//
//   -DLIMB_BITS=32                      32-bit limbs (three ~30-bit primes)
//   -DLIMB_BITS=64                      "hybrid": 32-bit limbs in and out; served
//                                       by the same 32-bit-limb product (exact, so
//                                       the limbs out are identical)
//   -DLIMB_BITS=64 -DNATIVE_HOST_LIMBS  64-bit limbs (three 58-bit primes, U160 CRT)
//
// As in the reference, `moduli` and `roots_of_unity_max` are declared extern.
// Define BW_DEFINE_TABLES in exactly one translation unit (before including
// this header) to emit their definitions for this configuration.
#pragma once
#ifndef CONFIG_H
#define CONFIG_H

#include <cstdint>
#include <type_traits>
#include <vector>

#ifndef LIMB_BITS
#define LIMB_BITS 32
#endif

#ifndef DEBUG
#define DEBUG 0
#endif

#if LIMB_BITS == 64 && defined(NATIVE_HOST_LIMBS)
#define NUM_MODULI 3
#define INPUT_LIMB_BITS 64
#define OUTPUT_LIMB_BITS 64
#define CRT_COEFF_BITS 160
#elif LIMB_BITS == 64
#define NUM_MODULI 2
#define INPUT_LIMB_BITS 32
#define OUTPUT_LIMB_BITS 32
#define CRT_COEFF_BITS 128
#else
#define NUM_MODULI 3
#define INPUT_LIMB_BITS 32
#define OUTPUT_LIMB_BITS 32
#define CRT_COEFF_BITS 128
#endif

#if LIMB_BITS == 64
using LimbType = uint64_t;
using TestDataTypeUint = uint64_t;
using TestDataTypeInt = int64_t;
#define LIMB_MASK 0xFFFFFFFFFFFFFFFFULL
#else
using LimbType = uint32_t;
using TestDataTypeUint = uint32_t;
using TestDataTypeInt = int32_t;
#define LIMB_MASK 0xFFFFFFFFULL
#endif

constexpr int BIT_WIDTH = sizeof(TestDataTypeUint) * 8;

using InputLimbType = std::conditional_t<INPUT_LIMB_BITS == 64, uint64_t, uint32_t>;
using OutputLimbType = std::conditional_t<OUTPUT_LIMB_BITS == 64, uint64_t, uint32_t>;

#if OUTPUT_LIMB_BITS == 64
#define OUTPUT_LIMB_MASK 0xFFFFFFFFFFFFFFFFULL
#else
#define OUTPUT_LIMB_MASK 0xFFFFFFFFULL
#endif

using namespace std;

#define BATCH 1
#define CARRY_SEG 1024

extern vector<TestDataTypeUint> moduli;
extern vector<TestDataTypeUint> roots_of_unity_max;

#ifdef BW_DEFINE_TABLES
// The reference's tables for this configuration (src/gpu_ntt.cu:28-47).
#if LIMB_BITS == 64 && defined(NATIVE_HOST_LIMBS)
vector<TestDataTypeUint> moduli = {0x400002600000001ULL, 0x400004200000001ULL, 0x400001100000001ULL};
vector<TestDataTypeUint> roots_of_unity_max = {273765203699653965ULL, 26231613454922890ULL, 144261359744151243ULL};
#elif LIMB_BITS == 64
vector<TestDataTypeUint> moduli = {0x400002600000001ULL, 0x400004200000001ULL};
vector<TestDataTypeUint> roots_of_unity_max = {273765203699653965ULL, 26231613454922890ULL};
#else
vector<TestDataTypeUint> moduli = {0x2d000001, 0x23800001, 0x26800001};
vector<TestDataTypeUint> roots_of_unity_max = {663, 721, 19};
#endif
#endif

// Digit width handed to libbirdwing_hip.so for this configuration.
#if LIMB_BITS == 64 && defined(NATIVE_HOST_LIMBS)
#define BW_CFG_LIMB_BITS 64
#else
#define BW_CFG_LIMB_BITS 32
#endif

#endif  // CONFIG_H
