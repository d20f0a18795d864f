"""MI355X-native batched Rodas3 chemistry step: a drop-in for the reference's CudaSolver call."""
