// Tile VM kernels with 2 rows per lane (see vm_device.h).
#include "vm_device.h"

namespace blz {
BLZ_VM_SHAPE_ENTRY(2)
}  // namespace blz
