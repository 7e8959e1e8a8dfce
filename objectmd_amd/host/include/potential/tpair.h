#ifndef OMD_POTENTIAL_TPAIR_H
#define OMD_POTENTIAL_TPAIR_H
#include <omd/system.h>
namespace omd {
// Tabulated pair potential pair.<name> (include/potential/tpair.h:38-72).
class TForcePair : public ForceKernel {
  string tablefile;

 public:
  explicit TForcePair(const string& TableFile) : tablefile(TableFile) {
    set_name("TABLE PAIR FORCE");
  }
  int Lower(omdg_t eng) override;
};
}  // namespace omd
#endif
