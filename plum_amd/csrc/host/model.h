// Host data model with the reference's calling surface (Bead / Molecule), so
// that Simulation-style callers of ForceField compile unchanged.  Method
// names follow molecules/bead.h:14-89 and molecules/molecule.h:15-68.
#pragma once
#include <string>
#include <vector>

namespace plum {

class Bead {
 public:
  Bead() = default;
  Bead(const std::string& symbol, int id, int chain_id, double q, double x, double y, double z)
      : symbol_(symbol), id_(id), chain_id_(chain_id), q_(q) {
    cur_[0] = trial_[0] = x;
    cur_[1] = trial_[1] = y;
    cur_[2] = trial_[2] = z;
  }
  double Charge() const { return q_; }
  void SetCharge(double q) { q_ = q; }
  int ID() const { return id_; }
  void SetID(int id) { id_ = id; }
  int ChainID() const { return chain_id_; }
  void SetChainID(int c) { chain_id_ = c; }
  const std::string& Symbol() const { return symbol_; }
  void SetSymbol(const std::string& s) { symbol_ = s; }
  int Type() const { return type_; }
  void SetType(int t) { type_ = t; }
  bool GetMoved() const { return moved_; }
  void SetMoved() { moved_ = true; }
  void UnsetMoved() { moved_ = false; }
  // flag 0: current position, 1: trial position (bead.cc:78-91)
  double GetCrd(int flag, int i) const { return flag == 0 ? cur_[i] : trial_[i]; }
  void SetCrd(int flag, int i, double v) {
    if (flag == 0) cur_[i] = v;
    else trial_[i] = v;
  }
  void SetAllCrd(const double p[3]) {
    for (int i = 0; i < 3; i++) cur_[i] = trial_[i] = p[i];
  }
  void UpdateCurrentPos() {
    for (int i = 0; i < 3; i++) cur_[i] = trial_[i];
  }
  void UpdateTrialPos() {
    for (int i = 0; i < 3; i++) trial_[i] = cur_[i];
  }

 private:
  double cur_[3] = {0, 0, 0};
  double trial_[3] = {0, 0, 0};
  std::string symbol_ = "NONE";
  int id_ = -1, chain_id_ = -1;
  int type_ = 0;
  double q_ = 0;
  bool moved_ = false;
};

class Molecule {
 public:
  std::vector<Bead> bds;
  std::vector<std::vector<int>> bonds;
  void AddBead(const Bead& b) { bds.push_back(b); }
  void AddBond(int a, int b) { bonds.push_back({a, b}); }
  int Size() const { return (int)bds.size(); }
};

}  // namespace plum
