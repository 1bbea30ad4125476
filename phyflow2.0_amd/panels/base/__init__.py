"""Mirror of pHyFlow.panels.base: panel operators."""
