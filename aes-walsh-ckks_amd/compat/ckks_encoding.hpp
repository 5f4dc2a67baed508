// Include-path shim for callers of the reference's headers (aes_bench.cpp
// includes ckks_encoding.hpp): the whole host API lives in host/ckks.hpp.
#pragma once
#include "../host/ckks.hpp"
