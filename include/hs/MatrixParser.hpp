// Text formats of the reference (src/utility/MatrixParser.cpp:11-274):
// "# N" header, then N lower-triangle rows of ';'-separated values; the
// right-hand side as "# N" and one ';'-separated row.
#pragma once

#include <string>
#include <vector>

#include "hs/Queue.hpp"
#include "hs/RightHandSide.hpp"
#include "hs/SymmetricMatrix.hpp"

class MatrixParser {
public:
    static SymmetricMatrix parseSymmetricMatrix(std::string& path, hs::Queue& queue);
    static RightHandSide parseRightHandSide(std::string& path, hs::Queue& queue);
    static std::vector<conf::fp_type> getRowValuesFromString(const std::string& rowString);
    // the full symmetric matrix, padding included, one row per line
    // (MatrixParser.cpp:214-274)
    static void writeFullMatrix(const std::string& path, const SymmetricMatrix& matrix);
    // the stored lower block triangle by block rows (MatrixParser.cpp:170-211)
    static void writeBlockedMatrix(const std::string& path, const SymmetricMatrix& matrix);
};
