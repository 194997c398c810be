// communicator.cpp -- the single-process Communicator (src/Communicator.cpp:14-213,
// src/communicators/CommGrid.cpp:186-345, CommUnstructured.cpp:87-149).
//
// One process owns every degree of freedom: the reference's commSize == 1
// branches.  This process is rank 0 of a world of one.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

#include "minim/Communicator.h"
#include "minim/Potential.h"

namespace minim {

int Communicator::getBlock(int loc) const {
  if (loc < 0 || (size_t)loc >= ndof) throw std::invalid_argument("Communicator: Invalid location");
  return 0;
}

int Communicator::getLocalIdx(int loc, int block) const {
  if (block == -1) block = getBlock(loc);
  return (commRank == block) ? loc : -1;
}

double Communicator::get(const vector<double>& v, int loc) const { return v[loc]; }

vector<double> Communicator::gather(const vector<double>& block, int root) const {
  if (!usesThisProc) return vector<double>();
  return block;
}

vector<double> Communicator::scatter(const vector<double>& data, int root) const {
  if (!usesThisProc) return vector<double>();
  return data;
}

double Communicator::sum(double a) const {
  if (!usesThisProc) return 0;
  return a;
}

double Communicator::sum(const vector<double>& a) const {
  if (!usesThisProc) return 0;
  double s = 0;
  for (double v : a) s += v;
  return sum(s);
}

double Communicator::norm(const vector<double>& a) const {
  if (!usesThisProc) return 0;
  return std::sqrt(dotProduct(a, a));
}

double Communicator::dotProduct(const vector<double>& a, const vector<double>& b) const {
  if (!usesThisProc) return 0;
  return sum(std::inner_product(a.begin(), a.end(), b.begin(), 0.0));
}

// Communicator::defaultSetup (src/Communicator.cpp:186-213) with a world of
// one process of rank 0.
void Communicator::defaultSetup(const Potential& pot, size_t n, vector<int> rk) {
  (void)pot;
  commSize = 1;
  commRank = 0;
  ndof = n;
  nproc = n;
  nblock = n;
  ranks = rk.empty() ? vector<int>{0} : rk;
  usesThisProc = false;
  for (int r : ranks)
    if (r == 0) usesThisProc = true;
  if (!usesThisProc) {
    commRank = -1;
    commSize = -1;
    nblock = (size_t)-1;
    nproc = (size_t)-1;
  }
}

void CommUnstructured::setup(Potential& pot, size_t n, vector<int> rk) { defaultSetup(pot, n, rk); }

// CommGrid::setup (CommGrid.cpp:291-345) with one process along every
// dimension: no halo, the local grid is the global one.
void CommGrid::setup(Potential& pot, size_t n, vector<int> rk) {
  defaultSetup(pot, n, rk);
  globalSizes = pot.gridSize;
  nDim = (int)pot.gridSize.size();
  if (!usesThisProc) {
    globalSizes = vector<int>(nDim + 1, 0);
    blockSizes = vector<int>(nDim + 1, 0);
    procSizes = vector<int>(nDim + 1, 0);
    procStart = vector<int>(nDim + 1, 0);
    haloWidths = vector<int>(nDim + 1, 0);
    commArray = vector<int>(nDim, 0);
    return;
  }
  if (commArray.empty()) commArray = pot.commArray;
  if (commArray.empty()) commArray = vector<int>(nDim, 1);
  if ((int)commArray.size() != nDim)
    throw std::invalid_argument("CommGrid: commArray has a size different to the grid dimensions.");
  int prod = 1;
  for (int c : commArray) prod *= c;
  if (prod != commSize)
    throw std::invalid_argument("CommGrid: The number of processors given by commArray does not match commSize.");
  commIndices = vector<int>(nDim, 0);
  haloWidths = vector<int>(nDim, 0);
  blockSizes = globalSizes;
  procSizes = globalSizes;
  procStart = vector<int>(nDim, 0);
  globalSizes.push_back(pot.dofPerNode);
  blockSizes.push_back(pot.dofPerNode);
  procSizes.push_back(pot.dofPerNode);
  haloWidths.push_back(0);
  procStart.push_back(0);
}

// CommGrid::dotProduct (CommGrid.cpp:186-194): every local site in order.
double CommGrid::dotProduct(const vector<double>& a, const vector<double>& b) const {
  if (!usesThisProc) return 0;
  double value = 0;
  const size_t n = std::min(a.size(), b.size());
  for (size_t i = 0; i < n; i++) value += a[i] * b[i];
  return sum(value);
}

}  // namespace minim
