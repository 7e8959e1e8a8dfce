#ifndef OMD_POTENTIAL_TEAM_H
#define OMD_POTENTIAL_TEAM_H
#include <omd/system.h>
namespace omd {
// Tabulated EAM: PAIR, EDENS and EMBED tables of eam.<name> found through
// $OMD_TABLE (src/gadget/team.cpp:206-222).  The embedding handler of the
// reference (TEmbedding) is part of the device kernel.
class TForceEAM : public ForceKernel {
  string tablefile;

 public:
  explicit TForceEAM(const string& PhiFile) : tablefile(PhiFile) { set_name("TABLE_FORCE_EAM"); }
  int Lower(omdg_t eng) override;
};
}  // namespace omd
#endif
