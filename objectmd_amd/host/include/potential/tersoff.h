#ifndef OMD_POTENTIAL_TERSOFF_H
#define OMD_POTENTIAL_TERSOFF_H
#include <omd/system.h>
namespace omd {
// Tersoff with the keys of tersoff.<material> (include/potential/tersoff.h:146-180)
// and its cutoff on [R-D, R+D] (tersoff.h:247-257).  The material may also be
// given by the system parameter tersoff.material, as the reference does.
class Tersoff : public ForceKernel {
  string material;

 public:
  explicit Tersoff(const string& mat = "") : material(mat) { set_name("TERSOFF"); }
  int Lower(omdg_t eng) override;
};
}  // namespace omd
#endif
