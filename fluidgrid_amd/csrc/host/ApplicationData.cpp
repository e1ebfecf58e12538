// ApplicationData.cpp -- the parameter singleton (ApplicationData.cpp:17-23).
#include "fluidgrid/ApplicationData.h"

namespace Simulation {
namespace Application {

ApplicationData* ApplicationData::instance() {
    static ApplicationData* inst = new ApplicationData();
    return inst;
}

}  // namespace Application
}  // namespace Simulation
