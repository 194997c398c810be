// Communicator.h -- the reference's Communicator API (include/Communicator.h:15-86,
// include/communicators/CommGrid.h:11-48, CommUnstructured.h:13-52) for one
// process driving one GPU.
//
// The reference distributes a State's degrees of freedom over MPI ranks; here
// one process owns the whole domain on its GPU, which is the reference's
// single-rank case (Communicator::defaultSetup, src/Communicator.cpp:186-213):
// commSize 1, ndof == nproc == nblock, halo exchanges and accumulations are
// no-ops, gather/scatter/assign are copies and the reductions are the local
// values.  A State constructed with a rank list that does not contain rank 0
// does not use this process (usesThisProc false, sizes -1), as in the
// reference.  Device vectors are reduced by the library's own FP64 kernels;
// these members serve host vectors, as the reference's do.
#ifndef MINIM_AMD_COMMUNICATOR_H
#define MINIM_AMD_COMMUNICATOR_H

#include <cstddef>
#include <memory>
#include <vector>

namespace minim {
using std::vector;
class Potential;

class Communicator {
 public:
  size_t ndof = 0;    // total number of degrees of freedom
  size_t nproc = 0;   // degrees of freedom on this process, including halo
  size_t nblock = 0;  // degrees of freedom assigned to this process

  bool usesThisProc = true;
  vector<int> ranks = vector<int>();

  int rank() const { return commRank; }
  int size() const { return commSize; }

  // Assign data: the local block (or processor range) of global data
  virtual vector<int> assignBlock(const vector<int>& in) const { return in; }
  virtual vector<char> assignBlock(const vector<char>& in) const { return in; }
  virtual vector<double> assignBlock(const vector<double>& in) const { return in; }
  virtual vector<int> assignProc(const vector<int>& in) const { return in; }
  virtual vector<char> assignProc(const vector<char>& in) const { return in; }
  virtual vector<double> assignProc(const vector<double>& in) const { return in; }

  // Access data
  virtual int getBlock(int loc) const;
  virtual int getLocalIdx(int loc, int block = -1) const;
  double get(const vector<double>& v, int loc) const;

  // Communication (Communicator.cpp:40-147)
  void communicate(vector<double>& v) const {}
  void communicateAccumulate(vector<double>& v) const {}
  vector<double> gather(const vector<double>& block, int root = -1) const;
  vector<double> scatter(const vector<double>& data, int root = -1) const;
  void bcast(int& value, int root = 0) const {}
  void bcast(double& value, int root = 0) const {}
  void bcast(vector<double>& value, int root = 0) const {}

  // Reductions (Communicator.cpp:150-173)
  double sum(double a) const;
  double sum(const vector<double>& a) const;
  double norm(const vector<double>& a) const;
  virtual double dotProduct(const vector<double>& a, const vector<double>& b) const;

  virtual ~Communicator() = default;
  virtual std::unique_ptr<Communicator> clone() const = 0;
  virtual void setup(Potential& pot, size_t ndof, vector<int> ranks) = 0;

 protected:
  int commSize = 1;
  int commRank = 0;
  vector<int> nGather;
  vector<int> iGather;
  void defaultSetup(const Potential& pot, size_t ndof, vector<int> ranks);
  virtual void makeMPITypes() {}
};

// Communicator for SERIAL and UNSTRUCTURED potentials (CommUnstructured.h:13-52).
class CommUnstructured : public Communicator {
 public:
  std::unique_ptr<Communicator> clone() const override { return std::make_unique<CommUnstructured>(*this); }
  void setup(Potential& pot, size_t ndof, vector<int> ranks) override;
};

// Communicator for GRID potentials (CommGrid.h:11-48): the whole grid on one
// process, no halo.
class CommGrid : public Communicator {
 public:
  explicit CommGrid(int haloWidth) : haloWidth(haloWidth) {}
  std::unique_ptr<Communicator> clone() const override { return std::make_unique<CommGrid>(*this); }
  void setup(Potential& pot, size_t ndof, vector<int> ranks) override;
  double dotProduct(const vector<double>& a, const vector<double>& b) const override;

  int nDim = 0;
  int haloWidth;
  vector<int> commArray;    // processes along each dimension (all 1)
  vector<int> commIndices;  // this process's indices (all 0)
  vector<int> globalSizes;  // grid sizes plus the DOFs per node (nDim + 1)
  vector<int> blockSizes;   // = globalSizes
  vector<int> procSizes;    // = globalSizes (no halo)
  vector<int> procStart;    // all 0
  vector<int> haloWidths;   // all 0
};

}  // namespace minim

#endif
