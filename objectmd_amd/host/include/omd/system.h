// Object-MD host API over the MI355X engine.
//
// The classes keep the names, signatures and call order that Object-MD user
// programs use (include/omd/system.h:139-410, include/omd/forcekernel.h:61-162,
// include/omd/container.h:76-248, include/omd/integrator.h:41-69), so a
// program written against the reference builds against these headers.  The
// difference is where the work happens: MDSystem::Run hands whole force
// evaluations and runs of steps to the device through the C ABI of omdg.h;
// Atoms(i) reads a host mirror that is refreshed from the device on demand.
#ifndef OMD_SYSTEM_H
#define OMD_SYSTEM_H

#include <cmath>
#include <fstream>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "omdg.h"

namespace omd {

using std::ofstream;
using std::ostream;
using std::string;
using std::vector;

#define NONPERIODIC 0
#define PERIODIC_X 1
#define PERIODIC_Y 2
#define PERIODIC_Z 4
#define PERIODIC 7

#define FLAG_ACTIVE 1
#define FLAG_OUTSIDE 2
#define FLAG_GHOST 4

#define MAXAUXVAR 16
#define MD_BOLTZMANN 8.61734279e-05   // eV/K (include/omd/conversion.h)
#define NORMAL_AMU 9.64853397731809e+3  // amu -> eV ps^2/A^2

// Per-atom record of the reference (include/omd/base.h:114-125).
struct Atom {
  int tid = 0, gid = 0, nid = 0, flag = FLAG_ACTIVE;
  double x = 0, y = 0, z = 0, vx = 0, vy = 0, vz = 0, fx = 0, fy = 0, fz = 0;
  double potential = 0, virial = 0;
  double aux[MAXAUXVAR] = {0};
};

struct SysBox {
  double x0 = 0, y0 = 0, z0 = 0, x1 = 0, y1 = 0, z1 = 0;
  double lx = 0, ly = 0, lz = 0, hlx = 0, hly = 0, hlz = 0;
};

// Errors: the reference throws const char* / std::string from die()
// (src/omd/omdtool.cpp:47-63); MDSystem::Run catches and returns failure.
[[noreturn]] void die(const string& msg);
void mdassert(bool cond, const string& msg);
void check_rc(int rc, const char* what);  // omdg status -> die(message)
string as_string(double v);

// Flat key/value parameter list (include/omd/param.h:41-101).
class ParamHandler {
 public:
  vector<string> spar;
  bool read(const string& filename);
  void append(const string& line);
  bool exist(const string& key) const;
  int index_of(const string& key) const;
  string string_value(const string& key, int idx = 0) const;
  double double_value(const string& key, int idx = 0) const;
  int int_value(const string& key, int idx = 0) const;
  bool peek(const string& key, double& v) const;
  bool peek(const string& key, int& v) const;
  bool peek(const string& key, string& v) const;
  bool peek(const string& key, bool& v) const;
};

// Table search path $OMD_TABLE (src/omd/base.cpp:36-45).
string search_path(const string& env, const string& file);

class MDSystem;

class MDGadget {
 protected:
  string name_;

 public:
  MDSystem* System = nullptr;
  ParamHandler* SysParam = nullptr;
  virtual ~MDGadget() {}
  virtual void Init(MDSystem* ws);
  virtual void ReadParameter() {}
  const string& get_name() const { return name_; }
  void set_name(const string& n) { name_ = n; }
  void register_class(const string&) {}
  virtual void PrintInfo(ostream& ost) { ost << name_ << "\n"; }
};

// Atom container: storage plus material (include/omd/container.h).
class AtomContainer : public MDGadget {
 protected:
  bool created = false;
  vector<Atom> atoms;

 public:
  string material;
  double M = 1.0;  // mass in eV ps^2/A^2 (container.cpp:94)
  double Z = 0.0;  // atomic number
  int id = 0;
  SysBox Box;
  ParamHandler param;

  explicit AtomContainer(const string& material_file = "");
  virtual ~AtomContainer() {}
  virtual AtomContainer* Create() {
    created = true;
    return this;
  }
  AtomContainer* SetName(const string& nm) {
    set_name(nm);
    return this;
  }
  AtomContainer* SetTemperature(double temperature);
  AtomContainer* SetKineticEnergy(double ek_per_atom);
  AtomContainer* SetVelocity(double vx, double vy, double vz);
  AtomContainer* Import(const string& filename);
  virtual SysBox& CalcBox();
  void Allocate(int n);
  int GetNAtom() const { return (int)atoms.size(); }
  Atom& Atoms(int i) { return atoms.at(i); }
  double GetMass(int) const { return M; }
  vector<Atom>& storage() { return atoms; }
  bool is_created() const { return created; }
  void set_id(int i) {
    id = i;
    for (auto& a : atoms) a.tid = i;
  }
};

// Force kernel plugin (include/omd/forcekernel.h:61-162).  Per-pair
// ComputeHalf/ComputeFull stay as the plugin authors' interface; the device
// engine needs a lowering, which the built-in kernels provide.
class ForceKernel : public MDGadget {
  friend class MDIntegrator;
  friend class MDSystem;

 protected:
  int AtomTypeA = 0, AtomTypeB = 0;
  double CutRadius = 0.0;

 public:
  virtual ~ForceKernel() {}
  ForceKernel* SetAtomID(int a, int b) {
    AtomTypeA = a;
    AtomTypeB = b;
    return this;
  }
  ForceKernel* SetCutRadius(double c) {
    CutRadius = c;
    return this;
  }
  bool CheckAtomID(int a, int b) const {
    return (a == AtomTypeA && b == AtomTypeB) || (a == AtomTypeB && b == AtomTypeA);
  }
  virtual void ComputeHalf(Atom&, Atom&) {}
  virtual void ComputeFull(Atom&, Atom&) {}
  // Device lowering: register tables/parameters with the engine and return
  // the potential kind (OMDG_POT_*).  Kernels without one cannot run here.
  virtual int Lower(omdg_t eng);
  double GetCutRadius() const { return CutRadius; }
};

// Modify hook classes (include/omd/modify.h).
class Modify : public MDGadget {};

// Iterator: the list radius parameters of VerletList (VerletList.cpp:46-52).
class MDIterator : public Modify {
 public:
  double CutRadius = -1.0;
  double RadiusTolerance = -1.0;
  bool full = false;
  void ReadParameter() override;
};

class Detector : public MDGadget {
 public:
  virtual void Detect() {}
  virtual int SampleEvery() const { return 1; }
};

class MDIntegrator : public MDGadget {
 public:
  double TimeStep = -1.0;
  vector<ForceKernel*> ActForces;
  explicit MDIntegrator(double dt = -1.0) : TimeStep(dt) { set_name("VERLET INTEGRATOR"); }
  ~MDIntegrator() override;
  MDIntegrator* AddForce(ForceKernel* f) {
    ActForces.push_back(f);
    return this;
  }
  void Init(MDSystem* ws) override;
};

// Main loop (include/omd/system.h:139-410).
class MDSystem : public AtomContainer {
 protected:
  vector<Atom> host;        // unified AoS mirror (UnificateAtoms order)
  bool host_stale = false;  // device is ahead of the mirror
  bool inline_default = false;
  omdg_t eng = nullptr;
  int argc_ = 0;
  char** argv_ = nullptr;

  virtual void CreateSystem() {}
  virtual void CreateGadget() = 0;
  virtual void SystemSetting() {}
  virtual void CreateGroup() {}
  virtual void BeforeRun() {}
  virtual void AfterRun() {}
  virtual void InlineFunction() { inline_default = true; }
  virtual bool CheckRun() { return ElapsedTime <= MaxTime; }

  void ReadParameter() override;
  void UnificateAtoms();
  void InitDevice();
  virtual void FirstRun();
  virtual void RunKernel();
  void SyncFromDevice();
  void PushToDevice();
  void Measure();

 public:
  MDIntegrator* Integrator = nullptr;
  MDIterator* Iterator = nullptr;
  vector<Modify*> Modifies;
  vector<Detector*> Detectors;
  vector<AtomContainer*> SystemAtoms;
  ParamHandler param;

  int PBoundary = 0;
  int Step = 0;
  double ElapsedTime = 0.0;
  double MaxTime = 0.0;
  double Energy = 0.0, Kinetic = 0.0, Virial = 0.0, Potential = 0.0, BasePotential = 0.0;
  int ExitCode = 0;
  bool silent_mode = true;

  MDSystem();
  ~MDSystem() override;

  void SetArgument(int argc, char* argv[]);
  int Run();

  AtomContainer* AddAtom(AtomContainer* a);
  MDSystem* AddForce(ForceKernel* f);
  MDSystem* AddForce(ForceKernel* f, AtomContainer* a, AtomContainer* b);
  MDSystem* AddForce(ForceKernel* f, int a, int b);
  MDSystem* AddModify(Modify* m);
  MDSystem* AddDetector(Detector* d);
  MDSystem* SetIntegrator(MDIntegrator* i);
  MDIterator* GetIterator() { return Iterator; }
  void SetOutputDirectory(const string&) {}

  Atom& Atoms(int i);
  int GetNAtom();
  void DumpAtoms(const string& filename);
  void PrintInfo(const string& filename);
  SysBox& CalcBox() override;
  omdg_t Engine() { return eng; }
};

// Periodic systems: the reference requires the grid class for PBC
// (VerletList.cpp:66-69); on one device it is the same loop.
class MDSystemGrid : public MDSystem {};

}  // namespace omd

#endif
