#ifndef OMD_CRYSTAL_DIAMOND_H
#define OMD_CRYSTAL_DIAMOND_H
#include <crystal/FCC.h>
namespace omd {
// Diamond("100", xml, yml, zml, material): xml/4+1 cells per axis
// (include/crystal/Diamond.h:22).
class Diamond : public Crystalline {
 public:
  Diamond(const string& orientation, int xml_, int yml_, int zml_, const string& mat_file)
      : Crystalline(orientation, xml_, yml_, zml_, mat_file) {
    set_name("DIAMOND");
  }
  AtomContainer* Create() override;
};
}  // namespace omd
#endif
