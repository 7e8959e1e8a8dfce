#ifndef OMD_DETECT_SYSMONITOR_H
#define OMD_DETECT_SYSMONITOR_H
#include <omd/system.h>
namespace omd {
// Energy monitor: "@ t: dt: Ek: Ep: E:" rows every monitor.every steps with
// Ep = Potential - BasePotential (include/detect/SysMonitor.h:35-134).
class SysMonitor : public Detector {
  string filename;
  std::ofstream fout;
  int every = 1;
  bool first = true;

 public:
  explicit SysMonitor(const string& fn = "md.out", bool = false) : filename(fn) {
    set_name("monitor");
  }
  void ReadParameter() override;
  void Init(MDSystem* ws) override;
  void Detect() override;
  int SampleEvery() const override { return every; }
};
}  // namespace omd
#endif
