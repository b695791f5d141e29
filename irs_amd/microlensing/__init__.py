"""Python interface of the engine, mirroring the reference's `microlensing`
package (microlensing/IPM, microlensing/Util) and adding microlensing.IRS."""
