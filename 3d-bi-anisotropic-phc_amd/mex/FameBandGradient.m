classdef FameBandGradient < handle
%FAMEBANDGRADIENT  Forward/backward pair for eigenvalue gradients over fame_mex.
%   The pencil is fncB x = omega fncA x (Lanczos' ew = 1/omega).  With a per-vertex
%   bi-isotropic material set by fame_mex('vertex_material', h, eps, gamma, mu, lattice)
%   on a plan with one k-point and N grid points:
%
%       g = FameBandGradient(h, N);
%       omega = g.forward(k, tol, maxiter);          % k smallest omega, ascending
%       [dL_deps, dL_dgamma] = g.backward(dL_domega); % N x 1 each
%
%   backward uses dw/dp = x' (d fncB/dp) x / x' fncA x (Hellmann-Feynman), with
%   x' (d fncB/dp) x per vertex from fame_mex('material_gradient', ...).  A cluster of
%   equal omega (relative gap <= cluster_rtol) gets the derivative of its mean for
%   every member; individual derivatives are not defined there.
%
%   Same algorithm as adjoint.BandGradient of the Python mirror (GPU-tested against
%   central finite differences); this file is source only, MATLAB being absent
%   where it was written.
    properties
        h                    % fame_mex plan handle
        N                    % grid points
        ev = []              % 4N x k eigenvectors of the forward pass, omega ascending
        omega = []           % k x 1
        cluster_rtol = 1e-8
    end
    methods
        function obj = FameBandGradient(h, N)
            obj.h = h;
            obj.N = N;
        end

        function omega = forward(obj, k, tol, maxiter)
            if nargin < 3, tol = 1e-10; end
            if nargin < 4, maxiter = 10000; end
            v0 = complex(ones(4 * obj.N, 1));
            [ev_, ew] = fame_mex('lanczos', obj.h, k, tol, maxiter, 2 * k, v0);
            [om, order] = sort(1 ./ ew(:), 'ascend');
            obj.ev = ev_(:, order);
            obj.omega = om;
            omega = om;
        end

        function [d_eps, d_gamma] = backward(obj, dL_domega)
            k = numel(obj.omega);
            g = fame_mex('material_gradient', obj.h, obj.ev);   % N x 3 x k
            den = zeros(1, k);
            for j = 1:k
                a = fame_mex('fncA', obj.h, obj.ev(:, j));
                den(j) = real(obj.ev(:, j)' * a);
            end
            G_eps = reshape(g(:, 1, :), obj.N, k) ./ den;    % N x k
            G_gam = reshape(g(:, 2, :), obj.N, k) ./ den;
            j = 1;
            while j <= k                                     % cluster means
                e = j;
                while e < k && abs(obj.omega(e + 1) - obj.omega(j)) <= obj.cluster_rtol * abs(obj.omega(j))
                    e = e + 1;
                end
                if e > j
                    G_eps(:, j:e) = repmat(mean(G_eps(:, j:e), 2), 1, e - j + 1);
                    G_gam(:, j:e) = repmat(mean(G_gam(:, j:e), 2), 1, e - j + 1);
                end
                j = e + 1;
            end
            w = dL_domega(:);
            d_eps = G_eps * w;
            d_gamma = G_gam * w;
        end
    end
end
