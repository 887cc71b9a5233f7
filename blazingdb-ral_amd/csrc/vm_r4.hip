// Tile VM kernels with 4 rows per lane (see vm_device.h).
#include "vm_device.h"

namespace blz {
BLZ_VM_SHAPE_ENTRY(4)
}  // namespace blz
