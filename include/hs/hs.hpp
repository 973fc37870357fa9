// Umbrella header of the C++ host API.
#pragma once
#include "hs/CG.hpp"
#include "hs/Cholesky.hpp"
#include "hs/Configuration.hpp"
#include "hs/LoadBalancer.hpp"
#include "hs/MatrixGenerator.hpp"
#include "hs/MatrixParser.hpp"
#include "hs/MetricsTracker.hpp"
#include "hs/Queue.hpp"
#include "hs/RightHandSide.hpp"
#include "hs/SymmetricMatrix.hpp"
#include "hs/TriangularSystemSolver.hpp"
