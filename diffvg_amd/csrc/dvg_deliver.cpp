// dvg_deliver.cpp -- hands out a render's gradient block as the .grad of its argument
// tensors (the direct leaf-gradient delivery of render_pytorch), one strided view per
// tensor, from C++.
//
// The autograd engine gives every argument tensor of RenderFunction its share of the
// scene gradients (render_pytorch.py:668-868: one gradient per argument, accumulated
// into .grad by the engine).  When those tensors are plain float32 leaves on the render
// device, the share of tensor i is the contiguous slice of the gradient block at
// kind_base[leaf_kind[i]] + leaf_off[i] with the tensor's shape.  Cutting tens of
// thousands of such views through Python costs microseconds each (a tensor object per
// view); here each view is made and stored in the tensor's autograd metadata without
// a Python object.  Host code only: no device work is launched except the in-place adds
// of tensors whose .grad already holds a value.
#include <torch/extension.h>

#include <vector>

namespace {

// For each i with want[i]: .grad = the block slice when .grad is undefined, else
// .grad += slice (in order, so a tensor that appears twice sums its slices).
void deliver(const at::Tensor &block, const std::vector<at::Tensor> &leaves, const std::vector<bool> &want,
             const at::Tensor &leaf_kind, const at::Tensor &leaf_off, const std::vector<int64_t> &kind_base) {
    TORCH_CHECK(block.dim() == 1 && block.is_contiguous() && block.scalar_type() == at::kFloat,
                "gradient block must be a contiguous float32 vector");
    TORCH_CHECK(leaf_kind.scalar_type() == at::kLong && leaf_off.scalar_type() == at::kLong &&
                    leaf_kind.is_contiguous() && leaf_off.is_contiguous() && !leaf_kind.is_cuda() &&
                    !leaf_off.is_cuda(),
                "leaf tables must be contiguous int64 host tensors");
    const int64_t n = (int64_t)leaves.size();
    TORCH_CHECK((int64_t)want.size() == n && leaf_kind.numel() == n && leaf_off.numel() == n,
                "leaf tables do not match the tensors");
    const int64_t *lk = leaf_kind.data_ptr<int64_t>();
    const int64_t *lo = leaf_off.data_ptr<int64_t>();
    const int64_t nkind = (int64_t)kind_base.size();
    const int64_t base = block.storage_offset(), len = block.numel();
    std::vector<int64_t> strides;
    for (int64_t i = 0; i < n; i++) {
        if (!want[i]) continue;
        const at::Tensor &t = leaves[i];
        TORCH_CHECK(lk[i] >= 0 && lk[i] < nkind, "leaf kind out of range");
        const at::IntArrayRef sizes = t.sizes();
        strides.assign(sizes.size(), 1);
        int64_t numel = 1;
        for (int64_t d = (int64_t)sizes.size() - 1; d >= 0; d--) {
            strides[d] = numel;
            numel *= sizes[d];
        }
        const int64_t off = kind_base[lk[i]] + lo[i];
        TORCH_CHECK(off >= 0 && off + numel <= len, "leaf slice outside the gradient block");
        at::Tensor g = block.as_strided(sizes, strides, base + off);
        at::Tensor &dst = const_cast<at::Tensor &>(t).mutable_grad();
        if (!dst.defined())
            dst = g;
        else
            dst.add_(g);
    }
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.def("deliver", &deliver, "gradient block slices into the .grad of the argument tensors");
}
