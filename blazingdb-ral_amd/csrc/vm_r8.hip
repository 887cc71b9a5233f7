// Tile VM kernels with 8 rows per lane (see vm_device.h).
#include "vm_device.h"

namespace blz {
BLZ_VM_SHAPE_ENTRY(8)
}  // namespace blz
