// MDSystem main loop over the device engine, and the lowering of the
// built-in force kernels onto the C ABI.
//
// Run() follows the reference's order (src/omd/system.cpp:152-177, 615-636,
// 860-873, 964-983): parameters, CreateSystem, CreateGadget, SystemSetting,
// atom unification, gadget initialisation, BeforeRun, FirstRun (one force
// evaluation, BasePotential, forces zeroed), then the step loop with
// InlineFunction and the detectors after every step, AfterRun.  Steps run on
// the device in batches as long as no host code needs the atoms in between:
// a batch ends where a detector samples or, when the program overrides
// InlineFunction, after every step.
#include <omd.h>

#include <cstdio>
#include <cstring>
#include <sstream>

namespace omd {

// ------------------------------------------------------------ lowering ----
// TForceEAM::Init + TEmbedding::AddTable (team.cpp:47-66, 206-222): the
// PAIR table of the type pair; for a species with itself its own EDENS and
// EMBED, for two species the cross EDENS, all from eam.<name>.
int TForceEAM::Lower(omdg_t eng) {
  const string path = search_path("$OMD_TABLE", "eam." + tablefile);
  const int a = AtomTypeA, b = AtomTypeB;
  check_rc(omdg_load_table(eng, OMDG_TAB_PAIR, a, b, path.c_str(), "PAIR"), "eam PAIR");
  check_rc(omdg_load_table(eng, OMDG_TAB_EDENS, a, b, path.c_str(), "EDENS"), "eam EDENS");
  if (a == b)
    check_rc(omdg_load_table(eng, OMDG_TAB_EMBED, a, a, path.c_str(), "EMBED"), "eam EMBED");
  return OMDG_POT_EAM;
}

int TForcePair::Lower(omdg_t eng) {
  const string path = search_path("$OMD_TABLE", "pair." + tablefile);
  check_rc(omdg_load_table(eng, OMDG_TAB_PAIR, AtomTypeA, AtomTypeB, path.c_str(), ""),
           "pair table");
  return OMDG_POT_PAIR;
}

int StillingerWeber::Lower(omdg_t eng) {
  ParamHandler p;
  if (!p.read(search_path("$OMD_TABLE", "sw." + paramfile)))
    die("can not find sw." + paramfile);
  bool usezbl = false;
  p.peek("zbl", usezbl);
  const double z1 = System->SystemAtoms.at(AtomTypeA)->Z;
  const double z2 = System->SystemAtoms.at(AtomTypeB)->Z;
  const double v[19] = {p.double_value("sw.epsilon"), p.double_value("sw.A"),
                        p.double_value("sw.B"),       p.double_value("sw.p"),
                        p.double_value("sw.q"),       p.double_value("sw.sigma"),
                        p.double_value("sw.a"),       p.double_value("sw.lambda"),
                        p.double_value("sw.gamma"),   p.double_value("sw.invers_cos0"),
                        usezbl ? 1.0 : 0.0,
                        usezbl ? p.double_value("spline.range", 0) : 0.0,
                        usezbl ? p.double_value("spline.range", 1) : 0.0,
                        usezbl ? p.double_value("spline.coeff", 0) : 0.0,
                        usezbl ? p.double_value("spline.coeff", 1) : 0.0,
                        usezbl ? p.double_value("spline.coeff", 2) : 0.0,
                        usezbl ? p.double_value("spline.coeff", 3) : 0.0,
                        z1,
                        z2};
  CutRadius = v[6] * v[5];
  check_rc(omdg_set_sw(eng, v), "sw parameters");
  return OMDG_POT_SW;
}

int Tersoff::Lower(omdg_t eng) {
  string mat = material;
  System->param.peek("tersoff.material", mat);
  ParamHandler p;
  if (!p.read(search_path("$OMD_TABLE", "tersoff." + mat))) die("can not find tersoff." + mat);
  const double v[13] = {p.double_value("tersoff.A"),        p.double_value("tersoff.B"),
                        p.double_value("tersoff.lambda_1"), p.double_value("tersoff.lambda_2"),
                        p.double_value("tersoff.lambda_3"), p.double_value("tersoff.alpha"),
                        p.double_value("tersoff.beta"),     p.double_value("tersoff.n"),
                        p.double_value("tersoff.c"),        p.double_value("tersoff.d"),
                        p.double_value("tersoff.h"),        p.double_value("tersoff.R"),
                        p.double_value("tersoff.D")};
  CutRadius = v[11] + v[12];  // tersoff.h: CutRadius = De + Re
  check_rc(omdg_set_tersoff(eng, v), "tersoff parameters");
  return OMDG_POT_TERSOFF;
}

void check_rc(int rc, const char* what) {
  if (rc != OMDG_OK) die(string(what) + ": " + omdg_last_error());
}

// ------------------------------------------------------------ the system ---
MDSystem::MDSystem() : AtomContainer("") { set_name("simulation_system"); }

MDSystem::~MDSystem() {
  if (eng) omdg_destroy(eng);
  delete Integrator;
  for (Modify* m : Modifies) delete m;
  for (Detector* d : Detectors) delete d;
  for (AtomContainer* a : SystemAtoms) delete a;
}

void MDSystem::SetArgument(int argc, char* argv[]) {
  argc_ = argc;
  argv_ = argv;
}

AtomContainer* MDSystem::AddAtom(AtomContainer* a) {
  a->set_id((int)SystemAtoms.size());
  a->System = this;
  SystemAtoms.push_back(a);
  return a;
}

MDSystem* MDSystem::SetIntegrator(MDIntegrator* i) {
  Integrator = i;
  return this;
}

MDSystem* MDSystem::AddForce(ForceKernel* f) {
  mdassert(Integrator != nullptr, "SetIntegrator before AddForce");
  Integrator->AddForce(f);
  return this;
}

MDSystem* MDSystem::AddForce(ForceKernel* f, int a, int b) {
  f->SetAtomID(a, b);
  return AddForce(f);
}

MDSystem* MDSystem::AddForce(ForceKernel* f, AtomContainer* a, AtomContainer* b) {
  return AddForce(f, a->id, b->id);
}

MDSystem* MDSystem::AddModify(Modify* m) {
  if (MDIterator* it = dynamic_cast<MDIterator*>(m)) Iterator = it;
  Modifies.push_back(m);
  return this;
}

MDSystem* MDSystem::AddDetector(Detector* d) {
  Detectors.push_back(d);
  return this;
}

// omd-parameter in the working directory (or --param <file>), then
// key value pairs from the command line (system.cpp:179-268).
void MDSystem::ReadParameter() {
  string file = "omd-parameter";
  for (int i = 1; i + 1 < argc_; i++)
    if (string(argv_[i]) == "--param") file = argv_[i + 1];
  param.read(file);
  for (int i = 1; i < argc_; i++) {
    if (string(argv_[i]) == "--param") {
      i++;
      continue;
    }
    param.append(argv_[i]);
  }
  param.peek("time.max", MaxTime);
  if (param.exist("boundary.periodic")) {
    string p = param.string_value("boundary.periodic");
    for (char& ch : p) ch = (char)std::tolower(ch);
    PBoundary = 0;
    if (p != "no" && p != "false") {
      if (p.find('x') != string::npos) PBoundary |= PERIODIC_X;
      if (p.find('y') != string::npos) PBoundary |= PERIODIC_Y;
      if (p.find('z') != string::npos) PBoundary |= PERIODIC_Z;
    }
  }
}

// All containers into one atom array, types = container index
// (system.cpp:413-444).
void MDSystem::UnificateAtoms() {
  host.clear();
  for (AtomContainer* c : SystemAtoms) {
    if (!c->is_created()) c->Create();
    for (const Atom& a : c->storage()) {
      Atom b = a;
      b.tid = c->id;
      b.nid = (int)host.size();
      host.push_back(b);
    }
  }
  mdassert(!host.empty(), "no atoms in the system");
}

SysBox& MDSystem::CalcBox() {
  SysBox& B = Box;
  bool first = true;
  for (AtomContainer* c : SystemAtoms) {
    bool inside = false;
    for (const Atom& a : c->storage()) inside |= !(a.flag & FLAG_OUTSIDE);
    if (!inside) continue;
    const SysBox& b = c->CalcBox();
    if (first) {
      B = b;
      first = false;
    } else {
      B.x0 = std::min(B.x0, b.x0);
      B.y0 = std::min(B.y0, b.y0);
      B.z0 = std::min(B.z0, b.z0);
      B.x1 = std::max(B.x1, b.x1);
      B.y1 = std::max(B.y1, b.y1);
      B.z1 = std::max(B.z1, b.z1);
    }
  }
  B.lx = std::fabs(B.x1 - B.x0);
  B.ly = std::fabs(B.y1 - B.y0);
  B.lz = std::fabs(B.z1 - B.z0);
  B.hlx = B.lx / 2;
  B.hly = B.ly / 2;
  B.hlz = B.lz / 2;
  return B;
}

void MDSystem::InitDevice() {
  check_rc(omdg_create(&eng, 0), "device");
  CalcBox();
  const double lo[3] = {Box.x0, Box.y0, Box.z0}, hi[3] = {Box.x1, Box.y1, Box.z1};
  check_rc(omdg_set_box(eng, lo, hi, PBoundary), "box");
  // potentials: lower every kernel; the combination must map to one device
  // potential (pair tables may accompany Stillinger-Weber for other types)
  int kind = OMDG_POT_NONE;
  double maxcut = 0.0;
  for (ForceKernel* f : Integrator->ActForces) {
    const int k = f->Lower(eng);
    if (kind == OMDG_POT_NONE || (kind == OMDG_POT_PAIR && k == OMDG_POT_SW))
      kind = k;
    else if (k != kind && !(kind == OMDG_POT_SW && k == OMDG_POT_PAIR))
      die("no device implementation for this combination of force kernels");
    maxcut = std::max(maxcut, f->GetCutRadius());
  }
  int fp64 = 0;
  if (kind == OMDG_POT_PAIR) {
    fp64 = 1;  // plain pair method in FP64 unless asked otherwise
    bool f32 = false;
    param.peek("device.fp32", f32);
    if (f32) fp64 = 0;
  }
  check_rc(omdg_set_potential(eng, kind, fp64), "potential");
  // list radius: cut + tolerance; tolerance from verlet.radtole or
  // comm.tolerance (absolute, or "%" of the cut), 10 % of the cut by default
  double skin = -1.0;
  if (Iterator && Iterator->RadiusTolerance > 0) skin = Iterator->RadiusTolerance;
  if (skin <= 0 && param.exist("comm.tolerance")) {
    string t = param.string_value("comm.tolerance");
    if (!t.empty() && t.back() == '%') skin = std::stod(t.substr(0, t.size() - 1)) * 0.01;
    else skin = std::stod(t);
    if (!t.empty() && t.back() == '%') skin *= (maxcut > 0 ? maxcut : 1.0);
  }
  if (skin <= 0) skin = 0.1 * (maxcut > 0 ? maxcut : 5.0);
  check_rc(omdg_set_neighbor(eng, skin, 10, 0), "neighbour list");
  PushToDevice();
}

void MDSystem::PushToDevice() {
  const int n = (int)host.size();
  vector<double> x(3 * (size_t)n), v(3 * (size_t)n);
  vector<int> tid(n), flag(n);
  for (int i = 0; i < n; i++) {
    const Atom& a = host[i];
    x[3 * i] = a.x;
    x[3 * i + 1] = a.y;
    x[3 * i + 2] = a.z;
    v[3 * i] = a.vx;
    v[3 * i + 1] = a.vy;
    v[3 * i + 2] = a.vz;
    tid[i] = a.tid;
    flag[i] = (a.flag & FLAG_ACTIVE) ? OMDG_FLAG_ACTIVE : 0;
  }
  vector<double> mass;
  for (AtomContainer* c : SystemAtoms) mass.push_back(c->M);
  check_rc(omdg_upload_atoms(eng, n, x.data(), v.data(), tid.data(), flag.data(),
                             (int)mass.size(), mass.data()),
           "atoms");
  host_stale = false;
}

void MDSystem::SyncFromDevice() {
  if (!host_stale) return;
  const int n = (int)host.size();
  vector<double> x(3 * (size_t)n), v(3 * (size_t)n), f(3 * (size_t)n), p(n);
  check_rc(omdg_download(eng, OMDG_X, x.data()), "download x");
  check_rc(omdg_download(eng, OMDG_V, v.data()), "download v");
  check_rc(omdg_download(eng, OMDG_F, f.data()), "download f");
  check_rc(omdg_download(eng, OMDG_POT, p.data()), "download potential");
  for (int i = 0; i < n; i++) {
    Atom& a = host[i];
    a.x = x[3 * i];
    a.y = x[3 * i + 1];
    a.z = x[3 * i + 2];
    a.vx = v[3 * i];
    a.vy = v[3 * i + 1];
    a.vz = v[3 * i + 2];
    a.fx = f[3 * i];
    a.fy = f[3 * i + 1];
    a.fz = f[3 * i + 2];
    a.potential = p[i];
  }
  host_stale = false;
}

Atom& MDSystem::Atoms(int i) {
  SyncFromDevice();
  return host.at(i);
}

int MDSystem::GetNAtom() { return (int)host.size(); }

void MDSystem::DumpAtoms(const string& filename) {
  SyncFromDevice();
  std::ofstream fl(filename);
  fl << "#$ Fields x y z vx vy vz --\n";
  char buf[256];
  for (const Atom& a : host) {
    std::snprintf(buf, sizeof buf, "%.17g %.17g %.17g %.17g %.17g %.17g\n", a.x, a.y, a.z,
                  a.vx, a.vy, a.vz);
    fl << buf;
  }
}

void MDSystem::PrintInfo(const string& filename) {
  std::ofstream fl(filename);
  fl << "atoms " << host.size() << " containers " << SystemAtoms.size() << "\n";
  fl << "box " << Box.x0 << " " << Box.x1 << " " << Box.y0 << " " << Box.y1 << " " << Box.z0
     << " " << Box.z1 << " periodic " << PBoundary << "\n";
  for (ForceKernel* f : Integrator->ActForces)
    fl << "force " << f->get_name() << " types " << f->AtomTypeA << "," << f->AtomTypeB
       << "\n";
  fl << "time.step " << Integrator->TimeStep << " time.max " << MaxTime << "\n";
}

void MDSystem::Measure() {
  Energy = Kinetic + Potential - BasePotential;
}

// One force evaluation; BasePotential = Ep; forces zeroed (system.cpp:860-873).
void MDSystem::FirstRun() {
  double ep = 0.0, vir = 0.0;
  check_rc(omdg_compute(eng, &ep, &vir), "first force evaluation");
  Potential = ep;
  Virial = vir;
  BasePotential = Potential;
  double ek = 0.0;
  for (const Atom& a : host) {
    const double m = SystemAtoms.at(a.tid)->M;
    if (a.flag & FLAG_ACTIVE) ek += m * (a.vx * a.vx + a.vy * a.vy + a.vz * a.vz);
  }
  Kinetic = 0.5 * ek;
  check_rc(omdg_zero_forces(eng), "zero forces");
  host_stale = true;
}

void MDSystem::RunKernel() {
  const double dt = Integrator->TimeStep;
  bool user_inline = true;
  bool probed = false;
  while (CheckRun()) {
    // steps until the next host-side sample
    int batch = 1;
    if (probed && !user_inline) {
      batch = 1 << 20;
      for (Detector* d : Detectors) {
        const int e = d->SampleEvery();
        batch = std::min(batch, (e - Step % e) % e + 1);
      }
    }
    // the loop condition ElapsedTime <= MaxTime with ElapsedTime += dt
    // (system.cpp:979, system.h:245), evaluated step by step
    double t = ElapsedTime;
    int k = 0;
    while (k < batch && t <= MaxTime) {
      t += dt;
      k++;
    }
    if (k == 0) break;
    double ek = 0.0, ep = 0.0;
    check_rc(omdg_step(eng, k, dt, &ek, &ep), "step");
    host_stale = true;
    Step += k - 1;
    for (int i = 0; i < k - 1; i++) ElapsedTime += dt;
    Kinetic = ek;
    Potential = ep;
    Measure();
    inline_default = false;
    InlineFunction();
    if (!probed) {
      user_inline = !inline_default;
      probed = true;
    }
    for (Detector* d : Detectors) d->Detect();
    ElapsedTime += dt;
    Step++;
  }
}

int MDSystem::Run() {
  try {
    ReadParameter();
    CreateSystem();
    CreateGadget();
    mdassert(Integrator != nullptr, "no integrator (SetIntegrator)");
    SystemSetting();
    param.peek("time.max", MaxTime);
    UnificateAtoms();
    CreateGroup();
    Integrator->Init(this);
    for (Modify* m : Modifies) m->Init(this);
    for (Detector* d : Detectors) d->Init(this);
    InitDevice();
    BeforeRun();
    FirstRun();
    RunKernel();
    AfterRun();
    ExitCode = 0;
  } catch (const string& e) {
    std::cerr << "error: " << e << std::endl;
    ExitCode = 1;
  } catch (const char* e) {
    std::cerr << "error: " << e << std::endl;
    ExitCode = 1;
  }
  return ExitCode;
}

}  // namespace omd
