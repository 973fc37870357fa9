// Padded right-hand side (src/dataStructures/RightHandSide.hpp:12-30):
// blockCountX * blockSize entries, zero past N.
#pragma once

#include <cstddef>
#include <vector>

#include "hs/Configuration.hpp"
#include "hs/Queue.hpp"

class RightHandSide {
public:
    RightHandSide(std::size_t N, int blockSize, hs::Queue& queue);

    const std::size_t N;
    const int blockSize;
    const int blockCountX;

    std::vector<conf::fp_type, hs::PinnedAllocator<conf::fp_type>> rightHandSideData;
};
