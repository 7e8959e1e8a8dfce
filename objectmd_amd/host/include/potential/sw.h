#ifndef OMD_POTENTIAL_SW_H
#define OMD_POTENTIAL_SW_H
#include <omd/system.h>
namespace omd {
// Stillinger-Weber with parameters from sw.<material> (src/gadget/sw.cpp:35-93).
class StillingerWeber : public ForceKernel {
  string paramfile;

 public:
  explicit StillingerWeber(const string& material) : paramfile(material) {
    set_name("STILLINGER WEBER");
  }
  int Lower(omdg_t eng) override;
};
}  // namespace omd
#endif
