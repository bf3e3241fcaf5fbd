// ppmg_ghost.hip -- ghost mapping of a padded field on the device, for one
// sub covering the domain (what ppm_topo_mkfield builds on one process).
//
// Replaces ppm_map_field_ghost with maptype ghost_get or ghost_put followed
// by push, send and pop (ppm_map_field_ghost.f:627-714):
//   ghost_get  every stored node that is a periodic image of a real node
//              (ghost layers, and the duplicate face node Nm) receives the
//              real node's value (pop replace);
//   ghost_put  every stored image node is added onto the real node it
//              images (pop add, ppm_map_field_pop_3d.f:781-800).
// The block lists are those of ppm_map_field_ghost_init.f:234-447 for one
// periodic sub: one block per periodic image of the sub (faces, edges,
// corners, in the reference's order), intersected with the stored box
// 1-g .. Nm+g (ppm_mesh_block_intersect.f:207-279).  As in the reference all
// blocks are packed before any is unpacked (push before pop), and blocks are
// unpacked in list order: a node in several image blocks keeps the last
// block's value (get) or receives the blocks' additions in list order (put).
//
// Element kinds are those of the reference's specifics (_s, _d, _sc, _dc, _i,
// _l): complex fields are real fields of twice the components; the push packs
// every kind into a real buffer and the pop converts back
// (ppm_map_field_push_3d.f:564-592, ppm_map_field_pop_3d.f:1818-1897), which
// for integers is an exact round trip (INT of the packed value) and for
// logicals 1.0 / 0.0 read back as value > 1 - eps, "add" being .AND.  The
// optional mask (one logical per stored node) is read at the pop: a node
// whose mask is .FALSE. keeps its value, and the buffer still steps past it
// (the push packs every node: its mask test is commented out at
// ppm_map_field_push_3d.f:546-549).
#include <type_traits>

#include "ppmg_internal.h"
#include "ppmg_interp.h"

namespace ppmg {

int ghost_blocks(const DevMesh& m, const int* g, GhostBlock* out) {
    const int dim = m.dim;
    // shift of one periodic image per axis: Nm - 1 nodes
    const int op[3] = {m.N[0], m.N[1], dim > 2 ? m.N[2] : 0};
    const bool px = m.periodic[0], py = m.periodic[1], pz = dim > 2 && m.periodic[2];
    int shifts[26][3];
    int n = 0;
    auto push = [&](int a, int b, int c) {
        shifts[n][0] = a;
        shifts[n][1] = b;
        shifts[n][2] = c;
        ++n;
    };
    // images in the order ppm_map_field_ghost_init.f:236-425 lists them: the x
    // faces, then for each y face its x edges, then for each z face the same
    // pattern in x and y
    auto plane = [&](int c) {
        if (px) {
            push(op[0], 0, c);
            push(-op[0], 0, c);
        }
        if (py) {
            push(0, op[1], c);
            if (px) {
                push(op[0], op[1], c);
                push(-op[0], op[1], c);
            }
            push(0, -op[1], c);
            if (px) {
                push(op[0], -op[1], c);
                push(-op[0], -op[1], c);
            }
        }
    };
    plane(0);
    if (pz) {
        push(0, 0, op[2]);
        if (px) {
            push(op[0], 0, op[2]);
            push(-op[0], 0, op[2]);
        }
        if (py) {
            push(0, op[1], op[2]);
            if (px) {
                push(op[0], op[1], op[2]);
                push(-op[0], op[1], op[2]);
            }
            push(0, -op[1], op[2]);
            if (px) {
                push(op[0], -op[1], op[2]);
                push(-op[0], -op[1], op[2]);
            }
        }
        push(0, 0, -op[2]);
        if (px) {
            push(op[0], 0, -op[2]);
            push(-op[0], 0, -op[2]);
        }
        if (py) {
            push(0, op[1], -op[2]);
            if (px) {
                push(op[0], op[1], -op[2]);
                push(-op[0], op[1], -op[2]);
            }
            push(0, -op[1], -op[2]);
            if (px) {
                push(op[0], -op[1], -op[2]);
                push(-op[0], -op[1], -op[2]);
            }
        }
    }
    // each image of the real nodes 1..Nm, clipped to the stored box 1-g..Nm+g
    int nb = 0;
    for (int q = 0; q < n; ++q) {
        GhostBlock b{};
        bool any = true;
        for (int d = 0; d < 3; ++d) {
            if (d >= dim) {
                b.src[d] = 1;
                b.size[d] = 1;
                b.off[d] = 0;
                continue;
            }
            const int nm = m.N[d] + 1;
            const int lo = (1 + shifts[q][d]) > (1 - g[d]) ? 1 + shifts[q][d] : 1 - g[d];
            const int hi = (nm + shifts[q][d]) < (nm + g[d]) ? nm + shifts[q][d] : nm + g[d];
            if (hi < lo) any = false;
            b.src[d] = lo - shifts[q][d];
            b.size[d] = hi - lo + 1;
            b.off[d] = shifts[q][d];
        }
        if (any) out[nb++] = b;
    }
    return nb;
}

// One block, one direction.  Node (i,j,k) (1-based, the reference's
// numbering) of the block is src + (i,j,k) - 1; its image adds off.
//   PACK_SRC: buf <- field(src block)        (ghost_get push)
//   PACK_IMG: buf <- field(image block)      (ghost_put push)
//   POP_SET:  field(image block) <- buf      (ghost_get pop, replace)
//   POP_ADD:  field(src block) += buf        (ghost_put pop, add)
// LOGIC: the elements are logicals (nonzero = .TRUE.), set as 1 / 0; the pop
// adds by .AND.  mask (nullptr: every node) holds one logical per stored node.
enum GhostOp { PACK_SRC = 0, PACK_IMG = 1, POP_SET = 2, POP_ADD = 3 };

template <typename T, bool LOGIC>
__global__ void __launch_bounds__(256) k_ghost_block(T* __restrict__ field, FieldGeom fg, GhostBlock b,
                                                     T* __restrict__ buf, int op, const int* __restrict__ mask) {
    const int lda = fg.lda;
    const int64_t n = (int64_t)b.size[0] * b.size[1] * b.size[2] * lda;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % lda);
        int64_t r = e / lda;
        const int i = (int)(r % b.size[0]);
        r /= b.size[0];
        const int j = (int)(r % b.size[1]);
        const int k = (int)(r / b.size[1]);
        const bool img = op == PACK_IMG || op == POP_SET;
        // 0-based node index, then the stored box
        const int nx = b.src[0] + i - 1 + (img ? b.off[0] : 0);
        const int ny = b.src[1] + j - 1 + (img ? b.off[1] : 0);
        const int nz = b.src[2] + k - 1 + (img ? b.off[2] : 0);
        const int64_t node = ((int64_t)(nz - fg.lo[2]) * fg.ext[1] + (ny - fg.lo[1])) * fg.ext[0] + (nx - fg.lo[0]);
        const int64_t at = node * lda + c;
        if (op == PACK_SRC || op == PACK_IMG) {
            buf[e] = field[at];
            continue;
        }
        if (mask && mask[node] == 0) continue;
        if constexpr (LOGIC) {
            const bool v = buf[e] != T(0);
            field[at] = (op == POP_SET ? v : (field[at] != T(0) && v)) ? T(1) : T(0);
        } else if constexpr (std::is_integral<T>::value) {
            // two's-complement wrap on overflow, as the hardware add
            field[at] = op == POP_SET ? buf[e] : (T)((uint32_t)field[at] + (uint32_t)buf[e]);
        } else {
            field[at] = op == POP_SET ? buf[e] : field[at] + buf[e];
        }
    }
}

template <typename T, bool LOGIC>
static int ghost_typed(T* field, const FieldGeom& fg, const GhostBlock* b, int nb, bool get, const int* mask,
                       DevBuf& scratch, hipStream_t s) {
    int64_t base[27];
    base[0] = 0;
    for (int q = 0; q < nb; ++q) base[q + 1] = base[q] + (int64_t)b[q].size[0] * b[q].size[1] * b[q].size[2] * fg.lda;
    if (base[nb] == 0) return 0;
    if (scratch.ensure(sizeof(T) * (size_t)base[nb])) return PPMG_ERROR_FATAL;
    T* buf = scratch.as<T>();
    auto grid = [](int64_t n) { return dim3((unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536)); };
    // push: every block is read before any is written
    for (int q = 0; q < nb; ++q)
        hipLaunchKernelGGL((k_ghost_block<T, LOGIC>), grid(base[q + 1] - base[q]), dim3(256), 0, s, field, fg, b[q],
                           buf + base[q], get ? (int)PACK_SRC : (int)PACK_IMG, (const int*)nullptr);
    // pop, in list order (a launch per block keeps the order of overlapping blocks)
    for (int q = 0; q < nb; ++q)
        hipLaunchKernelGGL((k_ghost_block<T, LOGIC>), grid(base[q + 1] - base[q]), dim3(256), 0, s, field, fg, b[q],
                           buf + base[q], get ? (int)POP_SET : (int)POP_ADD, mask);
    PPMG_HIP(hipGetLastError());
    return 0;
}

int map_field_ghost(const DevMesh& m, int kind, const FieldGeom& fg_in, void* field, const int* g, bool get,
                    const int* mask, DevBuf& scratch, hipStream_t s) {
    GhostBlock b[26];
    const int nb = ghost_blocks(m, g, b);
    FieldGeom fg = fg_in;
    switch (kind) {
    case PPMG_KIND_D: return ghost_typed<double, false>((double*)field, fg, b, nb, get, mask, scratch, s);
    case PPMG_KIND_S: return ghost_typed<float, false>((float*)field, fg, b, nb, get, mask, scratch, s);
    case PPMG_KIND_I: return ghost_typed<int, false>((int*)field, fg, b, nb, get, mask, scratch, s);
    case PPMG_KIND_L: return ghost_typed<int, true>((int*)field, fg, b, nb, get, mask, scratch, s);
    case PPMG_KIND_DC:
        fg.lda *= 2;   // (re, im) pairs: each part maps and adds as a real component
        return ghost_typed<double, false>((double*)field, fg, b, nb, get, mask, scratch, s);
    case PPMG_KIND_SC:
        fg.lda *= 2;
        return ghost_typed<float, false>((float*)field, fg, b, nb, get, mask, scratch, s);
    default:
        set_error(PPMG_ERROR_ERROR, PPMG_ERR_ARGUMENT, "unknown element kind");
        return PPMG_ERROR_ERROR;
    }
}

}  // namespace ppmg
