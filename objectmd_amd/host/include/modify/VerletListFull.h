#ifndef OMD_MODIFY_VERLETLISTFULL_H
#define OMD_MODIFY_VERLETLISTFULL_H
#include <omd/system.h>
namespace omd {
// Full Verlet list (src/gadget/VerletListFull.cpp), required by the
// three-body kernels (sw.cpp:65-67).
class VerletListFull : public MDIterator {
 public:
  VerletListFull() {
    set_name("VERLET LIST FULL NEIGHBOR");
    full = true;
  }
};
}  // namespace omd
#endif
