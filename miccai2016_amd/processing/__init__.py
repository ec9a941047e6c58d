"""Drop-in processing modules (reference: src/processing/) on the HIP C-ABI."""
