#ifndef OMD_CRYSTAL_PROJECTILE_H
#define OMD_CRYSTAL_PROJECTILE_H
#include <omd/system.h>
namespace omd {
// One atom aimed at (tx, ty) from height h with polar angle theta and
// azimuth phi (degrees) and kinetic energy (eV) (include/crystal/Projectile.h:22-49).
class Projectile : public AtomContainer {
  double tx_, ty_, h_, theta_, phi_, energy_;

 public:
  Projectile(double tx, double ty, double h, double theta, double phi, double energy,
             const string& mat)
      : AtomContainer(mat), tx_(tx), ty_(ty), h_(h), theta_(theta), phi_(phi), energy_(energy) {
    set_name("PROJECTILE");
  }
  AtomContainer* Create() override;
};
}  // namespace omd
#endif
