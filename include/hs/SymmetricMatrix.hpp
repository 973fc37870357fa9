// Packed lower-block-triangle SPD storage; same fields and layout as the
// reference (src/dataStructures/SymmetricMatrix.hpp:36-54): blocks of
// blockSize x blockSize enumerated column by column, each row major, diagonal
// blocks holding both triangles, zero padding past N.
#pragma once

#include <cstddef>
#include <vector>

#include "hs/Configuration.hpp"
#include "hs/Queue.hpp"

class SymmetricMatrix {
public:
    SymmetricMatrix(std::size_t N, int blockSize, hs::Queue& queue);

    const std::size_t N;
    const int blockSize;
    const int blockCountXY;

    std::vector<conf::fp_type, hs::PinnedAllocator<conf::fp_type>> matrixData;

    // packed ID of block (I, J), I >= J
    std::size_t blockID(int I, int J) const {
        return static_cast<std::size_t>(J) * blockCountXY - static_cast<std::size_t>(J) * (J - 1) / 2 + (I - J);
    }
};
