#ifndef OMD_MODIFY_VERLETLIST_H
#define OMD_MODIFY_VERLETLIST_H
#include <omd/system.h>
namespace omd {
// Half Verlet list (src/gadget/VerletList.cpp).  On the device the engine
// keeps one full list; the half list is the j > i part of it.
class VerletList : public MDIterator {
 public:
  VerletList() { set_name("VERLET LIST"); }
};
}  // namespace omd
#endif
