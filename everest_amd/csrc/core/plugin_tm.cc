// plugin_tm.cc -- the stock worker plugin (libTM.so): the table-driven
// gfx950 kernels behind the reference plugin entry point
// (system/templates/GPUWorker/interface.snp:1-7).
#include "everest/data.h"

extern "C" corelib::GPUWorker *getWorker(int gpu) { return new corelib::GPUWorkerTM(gpu); }
